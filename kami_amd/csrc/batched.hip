// Batched GEMM over problems of uneven size: one persistent launch, a device work queue,
// per-size-bucket tile variants chosen by an autotuner, optional hipGraph execution.
//
// Reference: batched_gemm_{1d,2d,3d}_double_mma (src/batched_gemm/KAMI/1d/
// batched_gemm_1d_double_mma.cu:59-218) runs one uniform-size problem per CTA at fixed
// strides (:71-74) with the partition picked per size by the sweep script
// (scripts/fp64_batched_GH200.sh:96-123).  Here:
//   * problems carry their own {m, n, k, lda, ldb, ldc, A, B, C} (the kami_problem array);
//   * every problem is cut into output tiles; a tile is one 64-byte queue entry holding
//     its own operand pointers, strides, extents and variant (WorkItem);
//   * a persistent grid (CUs x resident workgroups) pops a batch of up to kQueueBatch
//     entries per atomic add (fewer when the queue is short, so that every workgroup
//     gets work), in descending-cost order (longest first), until the queue is empty;
//     the next batch is fetched while the current one's tiles compute;
//   * the variant of a tile is the tile shape and wave partition of its size bucket,
//     picked by timing candidates on the device at plan creation (or from priors);
//   * one execution is [reset the queue counter] -> [persistent kernel], captured once
//     into a hipGraph and replayed.  The counter is cleared by a one-thread kernel,
//     never by a memset node.
// M, N and K must be multiples of 16.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "kami_internal.hpp"

namespace kami {

// One queue entry = one output tile, carrying everything its loads and stores need, so
// the kernel never goes back to the problem table.
struct WorkItem {
    const void* A;  // first element of the tile's rows of A
    const void* B;  // first element of the tile's rows of B (n x K) or columns (K x n)
    void* C;        // first element of the tile in C
    int64_t lda, ldb, ldc;
    int32_t mlim, nlim;  // rows / columns of C from the tile origin to the problem edge
    int32_t k, variant;
};
static_assert(sizeof(WorkItem) == 64, "queue entry layout");

// Queue entries popped per atomic add: a power of two from 2 to kQueueBatch (the LDS
// record size); the plan picks fewer for short queues.  Fetching a batch takes three
// tile boundaries (atomic add, entry loads, LDS write), so the next batch is requested
// when the workgroup starts the tile kFetchAhead before the end of its current batch,
// and never at the batch's first tile: a workgroup that starts early cannot take a
// second batch before the others have taken their first.
constexpr int kQueueShift = 3, kQueueBatch = 1 << kQueueShift;
constexpr int kFetchAhead = 3;

// Staging shared by every tile variant: a stage holds up to TMAX rows of A and TNMAX rows
// of B (or KC rows of TNMAX columns when B is K x N), KC elements of K each, in LDS rows
// padded to the conflict-free stride.  Pieces are 16 bytes; a thread owns pieces tid + i * NT.
template <class AT>
struct StreamGeom {
    static constexpr int NT = 256;
    static constexpr int ELT = (int)sizeof(typename AT::in_t), VE = 16 / ELT;
    // FP64 stages 8-wide K chunks of tiles up to 128 x 128 (96-byte rows, 48 KB of LDS
    // for the two stage buffers), FP16/BF16 64-wide chunks of tiles up to 128 x 64
    // (160-byte rows, 60 KB); two workgroups per CU either way.  FP64 offers the 128-wide
    // tiles because its batches are bound by HBM: a problem cut into 64 x 64 tiles reads
    // each row of A once per 64 columns of C and each row of B once per 64 rows, and
    // across the thousands of tiles in flight L2 keeps none of those rows for the next
    // tile that needs them
    static constexpr int KC = ELT == 8 ? 8 : 64;
    static constexpr int TMAX = 128, TNMAX = ELT == 8 ? 128 : 64;
    static constexpr int ROWB = lds_row_bytes(KC * ELT, AT::FRAG_BYTES);
    static constexpr int A_REGION = TMAX * ROWB;
    static constexpr int STAGE = (TMAX + TNMAX) * ROWB;
    static constexpr int LDS = 2 * STAGE;
    static constexpr int CPR = KC / VE;  // pieces per staged row
    static constexpr int NP = ((TMAX + TNMAX) * CPR + NT - 1) / NT;
    static_assert((KC * TNMAX) / VE <= TNMAX * CPR, "K x N staging fits the B region");
    static_assert((TMAX * CPR) % NT == 0 && (TNMAX * CPR) % NT == 0, "whole piece groups");
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
//   4: 128 x 128 tile, 2 x 2 waves (FP64)   5: 64 x 128 tile, 2 x 2 waves (FP64)
constexpr int kVariants = 6;
static_assert(kVariants == kBatchedVariants, "tuning-cache variant range");
static const int kVariantTM[kVariants] = {64, 32, 64, 128, 128, 64};
static const int kVariantTN[kVariants] = {64, 32, 64, 64, 128, 128};

template <class AT>
struct Variants {
    using V0 = TileGeom<AT, 64, 64, 2, 2>;
    using V1 = TileGeom<AT, 32, 32, 2, 2>;
    using V2 = TileGeom<AT, 64, 64, 4, 1>;
    using V3 = TileGeom<AT, 128, 64, 2, 2>;
    using V4 = TileGeom<AT, 128, 128, 2, 2>;
    using V5 = TileGeom<AT, 64, 128, 2, 2>;
};

// Variants a precision offers: FP16/BF16 stage 64 columns of B, so not 4 and 5.
inline int variants_of(int prec) { return prec == KAMI_PREC_F64 ? kVariants : 4; }

__device__ inline int variant_tm(int v) { return v == 1 ? 32 : (v == 3 || v == 4 ? 128 : 64); }
__device__ inline int variant_tn(int v) { return v == 1 ? 32 : (v == 4 || v == 5 ? 128 : 64); }

struct BatchArgs {
    const WorkItem* items;
    unsigned int* counter;
    int64_t n_items;
    double alpha, beta;
    int32_t b_kn;
    int32_t qshift;  // entries per pop = 1 << qshift, at most kQueueBatch
};

// A tile as the kernel holds it: wave-uniform values (in scalar registers).
struct Tile {
    const unsigned char* A;
    const unsigned char* B;
    unsigned char* C;
    int64_t lda, ldb, ldc;
    int mlim, nlim, k, variant, nch;
    bool valid;
};

__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ int64_t uni64(int64_t x) {
    const uint64_t u = (uint64_t)x;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)u);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(u >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}
template <class P>
__device__ __forceinline__ P* uniptr(P* p) {
    return reinterpret_cast<P*>(uni64((int64_t)reinterpret_cast<uintptr_t>(p)));
}

template <class AT>
__device__ __forceinline__ Tile read_tile(const WorkItem* rec, bool valid) {
    Tile t;
    t.valid = valid;
    if (!valid) {
        t.A = t.B = t.C = nullptr;
        t.lda = t.ldb = t.ldc = 0;
        t.mlim = t.nlim = t.k = t.variant = t.nch = 0;
        return t;
    }
    const WorkItem w = *rec;
    t.A = uniptr(reinterpret_cast<const unsigned char*>(w.A));
    t.B = uniptr(reinterpret_cast<const unsigned char*>(w.B));
    t.C = uniptr(reinterpret_cast<unsigned char*>(w.C));
    t.lda = uni64(w.lda);
    t.ldb = uni64(w.ldb);
    t.ldc = uni64(w.ldc);
    t.mlim = uni(w.mlim);
    t.nlim = uni(w.nlim);
    t.k = uni(w.k);
    t.variant = uni(w.variant);
    t.nch = (t.k + StreamGeom<AT>::KC - 1) / StreamGeom<AT>::KC;
    return t;
}

// Raw buffer descriptors over a tile's operands.  An offset at or past kBufferRange reads
// as zero, so pieces outside C or past K need no branch.
constexpr int kBufferRange = 0x7FFFFFF0;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t tile_rsrc(const unsigned char* base) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<unsigned char*>(base), (short)0,
                                             kBufferRange, 0x00020000);
}
__device__ __forceinline__ u4_t buffer_load16(__amdgpu_buffer_rsrc_t r, int off) {
    return __builtin_bit_cast(u4_t, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
}

// Loads the pieces of stage (tile t, K chunk c) into registers.  Every stage covers the
// full staging area (TMAX rows of A, TNMAX rows of B, or KC rows of TNMAX columns when B
// is K x N) with a fixed piece-to-thread mapping; rows outside the tile or C and K past
// the problem read as zero (no memory traffic), so the code is the same for every tile
// variant and free of branches.
template <class AT, bool KN>
__device__ __forceinline__ void stream_load(const Tile& t, int c, u4_t (&r)[StreamGeom<AT>::NP]) {
    using SG = StreamGeom<AT>;
    constexpr int CA = SG::TMAX * SG::CPR, CB = SG::TNMAX * SG::CPR;
    const int k0 = c * SG::KC;
    const int mrows = min(variant_tm(t.variant), t.mlim);
    const int ncols = min(variant_tn(t.variant), t.nlim);
    const __amdgpu_buffer_rsrc_t ra = tile_rsrc(t.A), rb = tile_rsrc(t.B);
    const int lda = (int)t.lda, ldb = (int)t.ldb;
#pragma unroll
    for (int i = 0; i < SG::NP; ++i) {
        const int idx = threadIdx.x + i * SG::NT;
        if (i * SG::NT < CA) {  // compile-time: the group is all A
            const int row = idx / SG::CPR, kc = (idx % SG::CPR) * SG::VE;
            const bool ok = row < mrows && k0 + kc < t.k;
            r[i] = buffer_load16(ra, ok ? (row * lda + k0 + kc) * SG::ELT : kBufferRange);
        } else if (i * SG::NT < CA + CB) {  // all B
            const int j = idx - CA;
            int off;
            if constexpr (!KN) {
                const int row = j / SG::CPR, kc = (j % SG::CPR) * SG::VE;
                off = (row < ncols && k0 + kc < t.k) ? (row * ldb + k0 + kc) * SG::ELT : kBufferRange;
            } else {
                constexpr int PPR = SG::TNMAX / SG::VE;  // pieces per K row
                const int kr = j / PPR, jc = (j % PPR) * SG::VE;
                off = (jc < ncols && k0 + kr < t.k) ? ((k0 + kr) * ldb + jc) * SG::ELT : kBufferRange;
            }
            r[i] = buffer_load16(rb, off);
        } else {
            r[i] = u4_t{0, 0, 0, 0};
        }
    }
}

template <class AT, bool KN>
__device__ __forceinline__ void stream_write(const u4_t (&r)[StreamGeom<AT>::NP], unsigned char* buf) {
    using SG = StreamGeom<AT>;
    using in_t = typename AT::in_t;
    constexpr int CA = SG::TMAX * SG::CPR, CB = SG::TNMAX * SG::CPR;
#pragma unroll
    for (int i = 0; i < SG::NP; ++i) {
        const int idx = threadIdx.x + i * SG::NT;
        if (i * SG::NT < CA) {
            store_raw<16>(buf + (idx / SG::CPR) * SG::ROWB + (idx % SG::CPR) * 16, r[i]);
        } else if (i * SG::NT < CA + CB) {
            const int j = idx - CA;
            if constexpr (!KN) {
                store_raw<16>(buf + SG::A_REGION + (j / SG::CPR) * SG::ROWB + (j % SG::CPR) * 16, r[i]);
            } else {
                constexpr int PPR = SG::TNMAX / SG::VE;
                const int kr = j / PPR, jc = (j % PPR) * SG::VE;
                const in_t* e = reinterpret_cast<const in_t*>(&r[i]);
#pragma unroll
                for (int q = 0; q < SG::VE; ++q)
                    *reinterpret_cast<in_t*>(buf + SG::A_REGION + (jc + q) * SG::ROWB + kr * SG::ELT) = e[q];
            }
        }
    }
}

// Shared state of a workgroup's stream of (tile, K chunk) stages.
//   * The queue is popped Q = qb = 1 << qs entries at a time into one of two LDS record
//     buffers; stream tile s lives in buffer (s / Q) % 2, slot s % Q.  The next batch
//     is fetched in steps taken at tile boundaries (atomic add; entry loads; LDS write),
//     so its latency hides behind whole tiles.
//   * Stage x of the stream (x counts (tile, chunk) pairs) is loaded into register slot
//     x % 2 and computed from LDS buffer x % 2.  While stage x is computed, slot
//     (x + 1) % 2 holds stage x + 1 and slot x % 2 holds stage x + 2.  The stage loop is
//     unrolled in pairs so both slot indices are compile-time, and it issues no memory
//     operation other than the stage loads: the compiler then waits only for the slot
//     it writes, never for the loads still in flight.
//   * The cursor (tile tc, chunk cc) is the next stage to load.  It stops ("blocked")
//     at a tile whose record has not been fetched yet; tile boundaries unblock it and
//     load any stage that is missing.
template <class AT>
struct Stream {
    Tile cur;            // tile being computed (stream index t0)
    Tile ct;             // tile at the cursor (stream index tc)
    int64_t t0, tc;
    int cc;              // chunk of the cursor
    bool blocked;
    int64_t x;           // stage being computed
    int64_t loaded;      // stages loaded so far
    int64_t fetched;     // stream tiles whose records are in LDS (whole batches)
    int qb, qs, lead;    // entries per batch (1 << qs); tile of a batch that requests the next
    int phase;           // next batch fetch: -1 idle, 0 start, 1 atomic issued,
                         // 2 entries loaded, 3 records written
    u4_t slot0[StreamGeom<AT>::NP], slot1[StreamGeom<AT>::NP];
    int sv0, sv1;        // variant of the stage in each slot, -1 = empty
    // wave 0 only: batch fetch in flight
    unsigned fbase;
    u4_t fpiece;         // lane l < 4 Q: 16-byte piece l % 4 of entry l / 4
    bool fvalid;
};

struct QueueLds {
    alignas(16) WorkItem rec[2][kQueueBatch];
    int valid[2][kQueueBatch];
};

template <class AT>
__device__ __forceinline__ Tile stream_tile(const QueueLds& q, int64_t s, int qs) {
    const int b = (int)((s >> qs) & 1), slot = (int)(s & ((1 << qs) - 1));
    return read_tile<AT>(&q.rec[b][slot], uni(q.valid[b][slot]) != 0);
}

// One step of the batch fetch (every thread, uniform); the caller puts a barrier after
// it and then calls fetch_commit.
template <class AT>
__device__ __forceinline__ void fetch_step(const BatchArgs& a, Stream<AT>& st, QueueLds& q) {
    st.phase = uni(st.phase);
    if (st.phase == 0) {
        // pop the queue range of stream tiles [fetched, fetched + Q)
        if (threadIdx.x == 0) st.fbase = atomicAdd(a.counter, (unsigned)st.qb);
        st.phase = 1;
    } else if (st.phase == 1) {
        if (threadIdx.x < 64) {
            const unsigned base = __shfl(st.fbase, 0);
            const int lane = threadIdx.x, e = lane >> 2;
            st.fvalid = e < st.qb && (int64_t)base + e < a.n_items;
            if (st.fvalid)
                st.fpiece = load_raw<16>(reinterpret_cast<const unsigned char*>(a.items + base + e) +
                                         16 * (lane & 3));
        }
        st.phase = 2;
    } else {
        const int b = (int)((st.fetched >> st.qs) & 1);
        if (threadIdx.x < 4 * st.qb) {
            const int e = threadIdx.x >> 2;
            if (st.fvalid)
                store_raw<16>(reinterpret_cast<unsigned char*>(&q.rec[b][e]) + 16 * (threadIdx.x & 3),
                              st.fpiece);
            if ((threadIdx.x & 3) == 0) q.valid[b][e] = st.fvalid ? 1 : 0;
        }
        st.phase = 3;
    }
}

template <class AT>
__device__ __forceinline__ void fetch_commit(Stream<AT>& st) {
    if (st.phase == 3) {
        st.fetched += st.qb;
        st.phase = -1;  // idle until the next batch is started
    }
}

// Makes the record of stream tile s available (uniform; synchronous).  The batch
// fetched here goes to the LDS buffer of the batch two before it, which the current
// tile and the cursor have both left.
template <class AT>
__device__ __forceinline__ void ensure_ready(const BatchArgs& a, Stream<AT>& st, QueueLds& q, int64_t s) {
    while (st.fetched <= s) {
        if (st.phase < 0) st.phase = 0;
        fetch_step(a, st, q);
        __syncthreads();
        fetch_commit(st);
    }
}

// Load the stage at the cursor into `dst` and advance the cursor.  Returns the stage's
// variant, or -1 when the cursor is at the end of the stream or blocked.  Only LDS is
// read besides the stage itself.
template <class AT, bool KN>
__device__ __forceinline__ int load_at_cursor(Stream<AT>& st, const QueueLds& q,
                                              u4_t (&dst)[StreamGeom<AT>::NP]) {
    if (!st.ct.valid || st.blocked) return -1;
    stream_load<AT, KN>(st.ct, st.cc, dst);
    const int v = st.ct.variant;
    ++st.loaded;
    if (++st.cc >= st.ct.nch) {
        st.cc = 0;
        ++st.tc;
        if (st.tc < st.fetched)
            st.ct = stream_tile<AT>(q, st.tc, st.qs);
        else
            st.blocked = true;
    }
    return v;
}

// Slot S holds the stage after the one being computed: write it to LDS buffer S and
// load the stage at the cursor into it.
template <class AT, bool KN, int S>
__device__ __forceinline__ void advance_slot(Stream<AT>& st, const QueueLds& q, unsigned char* smem) {
    using SG = StreamGeom<AT>;
    auto& slot = S ? st.slot1 : st.slot0;
    int& sv = S ? st.sv1 : st.sv0;
    if (sv >= 0) stream_write<AT, KN>(slot, smem + S * SG::STAGE);
    sv = load_at_cursor<AT, KN>(st, q, slot);
}

// Loads stage `loaded` into its slot (runtime parity; tile-boundary use only).
template <class AT, bool KN>
__device__ __forceinline__ void load_next_into_slot(Stream<AT>& st, const QueueLds& q) {
    if (st.loaded & 1)
        st.sv1 = load_at_cursor<AT, KN>(st, q, st.slot1);
    else
        st.sv0 = load_at_cursor<AT, KN>(st, q, st.slot0);
}

template <class AT, bool KN, class V>
__device__ __forceinline__ void run_tile(const BatchArgs& a, Stream<AT>& st, const QueueLds& q,
                                         unsigned char* smem) {
    using SG = StreamGeom<AT>;
    using acc_t = typename AT::acc_t;
    using out_t = typename AT::out_t;
    using frag_t = typename AT::frag_t;
    using accv_t = typename AT::accv_t;
    constexpr int FB = AT::FRAG_BYTES;
    const int lane = threadIdx.x & 63;
    const int wave = uni(threadIdx.x >> 6);
    const int wm = wave / V::WN, wn = wave % V::WN;
    const Tile cur = st.cur;
    // wave-uniform bounds
    const int mlim = uni(cur.mlim), nlim = uni(cur.nlim), nch = uni(cur.nch);

    accv_t acc[V::IM][V::JN];
#pragma unroll
    for (int i = 0; i < V::IM; ++i)
#pragma unroll
        for (int j = 0; j < V::JN; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[i][j][v] = 0;

    // Every atom of a live wave is computed: staged rows and columns outside C are zero,
    // and a guard per atom would make the compiler copy the accumulators around each
    // MFMA.  Waves whose whole sub-tile lies outside C skip the products.
    const bool wave_live = wm * V::SUBM < mlim && wn * V::SUBN < nlim;  // wave-uniform
    // atoms of the wave's sub-tile inside C (wave-uniform)
    const int im_live = uni(min(V::IM, (mlim - wm * V::SUBM + 15) / 16));
    const int jn_live = uni(min(V::JN, (nlim - wn * V::SUBN + 15) / 16));
    auto compute = [&](const unsigned char* base) {
        const unsigned char* bb = base + SG::A_REGION;
        if (!wave_live) return;
#pragma unroll
        for (int ks = 0; ks < SG::KC / AT::KSTEP; ++ks) {
            const int kk = (ks * AT::KSTEP + AT::KV * (lane >> 4)) * SG::ELT;
            frag_t bf[V::JN];
#pragma unroll
            for (int jn = 0; jn < V::JN; ++jn)
                bf[jn] = bitcast<frag_t>(load_raw<FB>(
                    bb + (wn * V::SUBN + jn * 16 + (lane & 15)) * SG::ROWB + kk));
            auto a_frag = [&](int im) {
                return bitcast<frag_t>(load_raw<FB>(
                    base + (wm * V::SUBM + im * 16 + (lane & 15)) * SG::ROWB + kk));
            };
            if constexpr (V::IM * V::JN >= 16) {
                // the 128 x 128 FP64 tile: its 128 accumulator VGPRs leave room for one
                // A fragment at a time; the JN products of a fragment's sub-step separate
                // the two sub-steps of one accumulator.  Row and column atoms past the
                // problem's edge are skipped (wave-uniform branches): M and N are
                // multiples of 16, so no product is wasted
#pragma unroll
                for (int im = 0; im < V::IM; ++im) {
                    if (im >= im_live) break;
                    const frag_t af = a_frag(im);
#pragma unroll
                    for (int e = 0; e < AT::NSUB; ++e)
#pragma unroll
                        for (int jn = 0; jn < V::JN; ++jn)
                            if (jn < jn_live) AT::mma_sub(acc[im][jn], af, bf[jn], e);
                }
            } else {
                frag_t af[V::IM];
#pragma unroll
                for (int im = 0; im < V::IM; ++im) af[im] = a_frag(im);
                // sub-steps of a fragment (FP64: two 16x16x4 instructions) interleaved
                // across the accumulators, so no instruction waits on the one just issued
#pragma unroll
                for (int e = 0; e < AT::NSUB; ++e)
#pragma unroll
                    for (int jn = 0; jn < V::JN; ++jn)
#pragma unroll
                        for (int im = 0; im < V::IM; ++im)
                            AT::mma_sub(acc[im][jn], af[im], bf[jn], e);
            }
        }
    };
    // stage of parity P: products from LDS buffer P, then the next stage (slot 1 - P) to
    // LDS buffer 1 - P and the cursor's stage into that slot
    auto stage0 = [&]() {
        compute(smem);
        advance_slot<AT, KN, 1>(st, q, smem);
        __syncthreads();
        ++st.x;
    };
    auto stage1 = [&]() {
        compute(smem + SG::STAGE);
        advance_slot<AT, KN, 0>(st, q, smem);
        __syncthreads();
        ++st.x;
    };
    int c = 0;
    if ((st.x & 1) && c < nch) {
        stage1();
        ++c;
    }
    for (; c + 1 < nch; c += 2) {
        stage0();
        stage1();
    }
    if (c < nch) stage0();

    const acc_t alpha = (acc_t)a.alpha, beta = (acc_t)a.beta;
    out_t* C = reinterpret_cast<out_t*>(cur.C);
#pragma unroll
    for (int im = 0; im < V::IM; ++im)
#pragma unroll
        for (int jn = 0; jn < V::JN; ++jn) {
            const int rb = wm * V::SUBM + im * 16, cb = wn * V::SUBN + jn * 16;
            if (rb >= mlim || cb >= nlim) continue;
            const int col = cb + (lane & 15);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = rb + AT::out_row(lane, v);
                out_t* dst = C + (int64_t)row * cur.ldc + col;
                acc_t val = alpha * acc[im][jn][v];
                if (beta != (acc_t)0) val += beta * (acc_t)(*dst);
                *dst = (out_t)val;
            }
        }
}

// Two workgroups per CU: the 128 x 128 FP64 tile holds 128 accumulator VGPRs per lane,
// and FP16/BF16 staging takes 60 KB of LDS.
template <class AT, bool KN>
__global__ void __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(2)))
batched_persistent_kernel(BatchArgs args) {
    using SG = StreamGeom<AT>;
    using VS = Variants<AT>;
    __shared__ __attribute__((aligned(16))) unsigned char smem[SG::LDS];
    __shared__ QueueLds q;
    Stream<AT> st;
    st.t0 = st.tc = 0;
    st.cc = 0;
    st.blocked = false;
    st.x = st.loaded = 0;
    st.fetched = 0;
    st.phase = -1;
    st.fvalid = false;
    st.sv0 = st.sv1 = -1;
    st.qs = min(kQueueShift, max(1, uni(args.qshift)));
    st.qb = 1 << st.qs;
    st.lead = max(1, st.qb - kFetchAhead);
    // prologue: the first batch synchronously
    ensure_ready(args, st, q, 0);
    st.cur = stream_tile<AT>(q, 0, st.qs);
    if (!st.cur.valid) return;  // uniform: the queue is empty
    st.ct = st.cur;
    // stage 0 into LDS buffer 0, stage 1 into slot 1, stage 2 into slot 0
    load_at_cursor<AT, KN>(st, q, st.slot0);
    stream_write<AT, KN>(st.slot0, smem);
    st.sv1 = load_at_cursor<AT, KN>(st, q, st.slot1);
    st.sv0 = load_at_cursor<AT, KN>(st, q, st.slot0);
    __syncthreads();
    while (true) {
        switch (st.cur.variant) {
            case 0: run_tile<AT, KN, typename VS::V0>(args, st, q, smem); break;
            case 1: run_tile<AT, KN, typename VS::V1>(args, st, q, smem); break;
            case 2: run_tile<AT, KN, typename VS::V2>(args, st, q, smem); break;
            case 3: run_tile<AT, KN, typename VS::V3>(args, st, q, smem); break;
            default:  // 4 and 5 exist where B is staged 128 columns wide
                if constexpr (SG::TNMAX >= 128) {
                    if (st.cur.variant == 4)
                        run_tile<AT, KN, typename VS::V4>(args, st, q, smem);
                    else
                        run_tile<AT, KN, typename VS::V5>(args, st, q, smem);
                }
                break;
        }
        // ---- tile boundary ----
        ++st.t0;
        // the next batch is requested at tile `lead` (at least 1) of the current one,
        // which is then fetched; a fetch in flight takes one step per tile boundary (the
        // buffer it fills is the one of the batch before the current one, which the
        // tile and the cursor have left)
        if (st.phase < 0 && (int)(st.t0 & (st.qb - 1)) == st.lead &&
            st.fetched == st.t0 - st.lead + st.qb)
            st.phase = 0;
        if (st.phase >= 0) {
            fetch_step(args, st, q);
            __syncthreads();
            fetch_commit(st);
        }
        ensure_ready(args, st, q, st.t0);  // uniform; normally already done
        st.cur = stream_tile<AT>(q, st.t0, st.qs);
        if (!st.cur.valid) break;  // uniform
        // a blocked cursor resumes; stages that could not be loaded are loaded now
        if (st.blocked) {
            ensure_ready(args, st, q, st.tc);
            st.ct = stream_tile<AT>(q, st.tc, st.qs);
            st.blocked = false;
            if (st.loaded == st.x) {
                // the new tile's first stage: through its slot straight to its LDS buffer
                load_next_into_slot<AT, KN>(st, q);
                if (st.x & 1) {
                    stream_write<AT, KN>(st.slot1, smem + SG::STAGE);
                    st.sv1 = -1;
                } else {
                    stream_write<AT, KN>(st.slot0, smem);
                    st.sv0 = -1;
                }
            }
            // then stages x + 1 and x + 2 into their slots, as far as the fetched tiles
            // reach: a cursor that blocks again (the batch just fetched holds fewer
            // stages than that) stays blocked until the next tile boundary repairs it
            while (st.loaded < st.x + 3 && st.ct.valid && !st.blocked)
                load_next_into_slot<AT, KN>(st, q);
            __syncthreads();
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
// (scripts/fp64_batched_GH200.sh:7-14).  FP64 takes 128 rows or columns wherever the
// problem has more than 64 of them (its batches are bound by HBM, and the wider tile
// halves the re-reads of the other operand).
static int prior_variant(int prec, int bucket) {
    const int cm = bucket / 25, cn = (bucket / 5) % 5;
    if (cm == 0 && cn == 0) return 1;
    if (prec == KAMI_PREC_F64) {
        if (cm >= 2 && cn >= 2) return 4;
        if (cm >= 2) return 3;
        if (cn >= 2) return 5;
        return 0;
    }
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

template <class AT>
static hipError_t launch_for(const BatchArgs& a, int grid, hipStream_t s) {
    if (a.b_kn)
        hipLaunchKernelGGL((batched_persistent_kernel<AT, true>), dim3(grid), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL((batched_persistent_kernel<AT, false>), dim3(grid), dim3(256), 0, s, a);
    return hipGetLastError();
}

// log2 of the entries per pop: kQueueBatch when every workgroup gets three batches or
// more, fewer for a short queue (4096 tiles over 768 workgroups: two entries per pop).
static int32_t queue_shift(int64_t n_items, int grid) {
    const int64_t per = n_items / (3 * (int64_t)std::max(grid, 1));
    int32_t sh = 1;
    while (sh < kQueueShift && (int64_t(2) << sh) <= per) ++sh;
    return sh;
}

static hipError_t launch_batched(int prec, const BatchArgs& a, int grid, hipStream_t s) {
    hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(64), 0, s, a.counter);
    if (prec == KAMI_PREC_F64) return launch_for<AtomF64>(a, grid, s);
    if (prec == KAMI_PREC_F16_F32) return launch_for<AtomF16x32>(a, grid, s);
    return launch_for<AtomBF16x32>(a, grid, s);
}

template <class AT>
static hipError_t occupancy_for(bool kn, int* per_cu) {
    if (kn)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, batched_persistent_kernel<AT, true>, 256, 0);
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, batched_persistent_kernel<AT, false>, 256, 0);
}

static hipError_t persistent_grid(int prec, bool kn, int* grid) {
    int dev = 0, cus = 0, per_cu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    if (prec == KAMI_PREC_F64)
        e = occupancy_for<AtomF64>(kn, &per_cu);
    else if (prec == KAMI_PREC_F16_F32)
        e = occupancy_for<AtomF16x32>(kn, &per_cu);
    else
        e = occupancy_for<AtomBF16x32>(kn, &per_cu);
    if (e != hipSuccess) return e;
    *grid = std::max(1, cus * std::max(1, per_cu));
    return hipSuccess;
}

// Queue entries of the problems: every problem cut into tiles of its bucket's variant,
// each entry carrying the tile's own operand pointers.  Problems are ordered longest
// first (by m n k) so the tail of the queue holds the cheapest work, and a problem's tiles
// stay consecutive in row-major order: a workgroup pops a batch of neighbouring tiles at
// once, so tiles that share rows of A or columns of B run on one CU, back to back, and
// find those operands in its L2 instead of fetching them again (ordering by tile cost
// scattered them over all XCDs and read A and B 1.8 times over).
// Queue entries of `count` problems (bucket[p] their size bucket), written to out[];
// `order` is room for `count` indices.  Equal-cost problems keep their input order
// (an index tie-break instead of a stable sort, which may allocate).
static int64_t count_items(const kami_problem* probs, int64_t count, const int* bucket,
                           const int* variant_of_bucket) {
    int64_t n = 0;
    for (int64_t p = 0; p < count; ++p) {
        const int v = variant_of_bucket[bucket[p]];
        n += (int64_t)((probs[p].m + kVariantTM[v] - 1) / kVariantTM[v]) *
             ((probs[p].n + kVariantTN[v] - 1) / kVariantTN[v]);
    }
    return n;
}

static void build_items(const kami_problem* probs, int64_t count, const int* bucket,
                        const int* variant_of_bucket, int ei, int eo, bool b_kn, int64_t* order,
                        WorkItem* out) {
    for (int64_t p = 0; p < count; ++p) order[p] = p;
    std::sort(order, order + count, [&](int64_t a, int64_t b) {
        const kami_problem &P = probs[a], &Q = probs[b];
        const int64_t ca = (int64_t)P.m * P.n * P.k, cb = (int64_t)Q.m * Q.n * Q.k;
        return ca != cb ? ca > cb : a < b;
    });
    int64_t i = 0;
    for (int64_t q = 0; q < count; ++q) {
        const int64_t p = order[q];
        const kami_problem& P = probs[p];
        const int v = variant_of_bucket[bucket[p]];
        const int tm = kVariantTM[v], tn = kVariantTN[v];
        for (int r = 0; r < P.m; r += tm)
            for (int c = 0; c < P.n; c += tn) {
                WorkItem& w = out[i++];
                w.A = static_cast<const char*>(P.A) + (int64_t)r * P.lda * ei;
                w.B = static_cast<const char*>(P.B) + (b_kn ? (int64_t)c : (int64_t)c * P.ldb) * ei;
                w.C = static_cast<char*>(P.C) + ((int64_t)r * P.ldc + c) * eo;
                w.lda = P.lda;
                w.ldb = P.ldb;
                w.ldc = P.ldc;
                w.mlim = P.m - r;
                w.nlim = P.n - c;
                w.k = P.k;
                w.variant = v;
            }
    }
}

static void build_items(const std::vector<kami_problem>& probs, const std::vector<int>& bucket,
                        const int* variant_of_bucket, int ei, int eo, bool b_kn,
                        std::vector<WorkItem>& items) {
    const int64_t count = (int64_t)probs.size();
    std::vector<int64_t> order(probs.size());
    items.resize((size_t)count_items(probs.data(), count, bucket.data(), variant_of_bucket));
    build_items(probs.data(), count, bucket.data(), variant_of_bucket, ei, eo, b_kn, order.data(),
                items.data());
}

// Validates one problem of a batch and fills in packed leading dimensions.
static kami_status check_problem(kami_problem& p, int b_layout, int ei, int eo) {
    if (p.m <= 0 || p.n <= 0 || p.k <= 0 || p.m % 16 || p.n % 16 || p.k % 16) return KAMI_EINVAL;
    if (!p.A || !p.B || !p.C) return KAMI_EINVAL;
    if (!p.lda) p.lda = p.k;
    if (!p.ldb) p.ldb = b_layout == KAMI_B_NK ? p.k : p.n;
    if (!p.ldc) p.ldc = p.n;
    if (p.lda < p.k || p.ldb < (b_layout == KAMI_B_NK ? p.k : p.n) || p.ldc < p.n)
        return KAMI_EINVAL;
    if (!span_fits(p.m, p.lda, ei) || !span_fits(b_layout == KAMI_B_NK ? p.n : p.k, p.ldb, ei) ||
        !span_fits(p.m, p.ldc, eo))
        return KAMI_EINVAL;
    if ((p.lda * ei) % 16 || (p.ldb * ei) % 16 || ((uintptr_t)p.A % 16) || ((uintptr_t)p.B % 16) ||
        ((uintptr_t)p.C % eo))
        return KAMI_EINVAL;
    // Operand pieces are addressed by 32-bit byte offsets from the tile's origin
    // (stream_load): the farthest piece of any tile must end below kBufferRange.
    const int64_t a_span = ((int64_t)(std::min(p.m, 128) - 1) * p.lda + p.k) * ei;
    const int64_t b_span = b_layout == KAMI_B_NK
                               ? ((int64_t)(std::min(p.n, 128) - 1) * p.ldb + p.k) * ei
                               : ((int64_t)(p.k - 1) * p.ldb + std::min(p.n, 128)) * ei;
    if (a_span > kBufferRange || b_span > kBufferRange) return KAMI_ENOTSUP;
    return KAMI_OK;
}

// Tile variant of every size bucket without timing: the tuning cache's choice, else the
// prior.
static void untimed_variants(int precision, int (&variant_of_bucket)[125]) {
    for (int b = 0; b < 125; ++b) {
        const int tuned = batched_tuned_variant(precision, b);
        variant_of_bucket[b] =
            tuned >= 0 && tuned < variants_of(precision) ? tuned : prior_variant(precision, b);
    }
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

// ---- one-shot execution without allocation ----
// The queue and its counter of kami_batched_gemm_static live in `scratch`, a file-scope
// device array (each device holds its own copy until the process ends): the counter in its
// first 256 bytes, then up to kStaticItems queue entries.  Host code reaches a device's
// copy only through the address hipGetSymbolAddress returns with that device current (the
// array's name is a host-side stand-in, not a device address).  The host side stages the
// problems, their order and the entries in static arrays, so the call makes no allocation
// of device or host memory; a mutex keeps two host threads from sharing the staging.
constexpr int64_t kStaticItems = 32768;  // 2 MiB of queue entries
constexpr size_t kStaticCounterBytes = 256;
static __device__ __align__(256) unsigned char scratch[kStaticCounterBytes + kStaticItems * sizeof(WorkItem)];

struct StaticStaging {
    std::mutex mu;
    unsigned char* dev_ptr[kMaxCachedDevices] = {};  // scratch of each device
    kami_problem probs[kStaticItems];  // a problem holds at least one entry
    int bucket[kStaticItems];
    int64_t order[kStaticItems];
    WorkItem items[kStaticItems];
};
static StaticStaging g_static;

// scratch of the current device (caller holds g_static.mu)
static kami_status static_scratch(unsigned char** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e);
    if (dev < 0 || dev >= kMaxCachedDevices) return KAMI_ENOTSUP;
    if (!g_static.dev_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e);
        g_static.dev_ptr[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_static.dev_ptr[dev];
    return KAMI_OK;
}

static void free_plan(kami_batched_plan* p) {
    if (!p) return;
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    if (p->cap_stream) (void)hipStreamDestroy(p->cap_stream);
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
        const kami_status pst = check_problem(p, o.b_layout, ei, eo);
        if (pst != KAMI_OK) return pst;
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
    untimed_variants(precision, variant_of_bucket);

    hipError_t e = persistent_grid(precision, plan->b_kn, &plan->grid);
    if (e != hipSuccess) {
        free_plan(plan);
        return hip_fail(e);
    }
    st = upload(nullptr, 0, (void**)&plan->d_counter);
    if (st != KAMI_OK) {
        free_plan(plan);
        return st;
    }

    if (o.autotune) {
        // time every variant on up to 256 problems of each bucket, keep the fastest.  The
        // sample is repeated until it carries at least kTuneProblems problems so that every
        // variant is timed at full occupancy (a small sample would be latency-bound and
        // favour the smallest tiles); the repeats write identical values to the same C
        // (tuning requires beta = 0).
        constexpr size_t kTuneProblems = 4096;
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
            const size_t distinct = sample.size();
            for (size_t i = 0; sample.size() < kTuneProblems; ++i) {
                sample.push_back(sample[i % distinct]);
                sb.push_back(0);
            }
            float best = 1e30f;
            for (int v = 0; v < variants_of(precision); ++v) {
                int vb[1] = {v};
                std::vector<WorkItem> items;
                build_items(sample, sb, vb, ei, eo, plan->b_kn, items);
                WorkItem* d_it = nullptr;
                st = upload(items.data(), items.size() * sizeof(WorkItem), (void**)&d_it);
                if (st != KAMI_OK) break;
                BatchArgs a{d_it, plan->d_counter, (int64_t)items.size(), o.alpha, 0.0, plan->b_kn,
                            queue_shift((int64_t)items.size(), plan->grid)};
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
    build_items(hp, bucket, variant_of_bucket, ei, eo, plan->b_kn, items);
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
            BatchArgs a{plan->d_items, plan->d_counter, plan->n_items, plan->alpha,
                        plan->beta, plan->b_kn, queue_shift(plan->n_items, plan->grid)};
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
        BatchArgs a{plan->d_items, plan->d_counter, plan->n_items, plan->alpha,
                    plan->beta, plan->b_kn, queue_shift(plan->n_items, plan->grid)};
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

kami_status kami_batched_gemm_static(const kami_problem* probs, int64_t count, int32_t precision,
                                     const kami_batched_opts* opts, void* stream) {
    if (!probs || count <= 0) return KAMI_EINVAL;
    if (precision != KAMI_PREC_F64 && precision != KAMI_PREC_F16_F32 &&
        precision != KAMI_PREC_BF16_F32)
        return KAMI_ENOTSUP;
    kami_batched_opts o;
    std::memset(&o, 0, sizeof(o));
    o.alpha = 1.0;
    if (opts) o = *opts;
    if (o.b_layout != KAMI_B_NK && o.b_layout != KAMI_B_KN) return KAMI_EINVAL;
    if (o.autotune || o.use_graph) return KAMI_EINVAL;  // both need memory of their own
    if (count > kStaticItems) return KAMI_ENOSPC;
    const int ei = elt_in(precision), eo = elt_out(precision);
    const bool b_kn = o.b_layout == KAMI_B_KN;

    std::lock_guard<std::mutex> lock(g_static.mu);
    int variant_of_bucket[125];
    untimed_variants(precision, variant_of_bucket);
    for (int64_t i = 0; i < count; ++i) {
        kami_problem& p = g_static.probs[i];
        p = probs[i];
        const kami_status pst = check_problem(p, o.b_layout, ei, eo);
        if (pst != KAMI_OK) return pst;
        g_static.bucket[i] = bucket_of(p);
    }
    const int64_t n_items = count_items(g_static.probs, count, g_static.bucket, variant_of_bucket);
    if (n_items > kStaticItems) return KAMI_ENOSPC;
    kami_status st = device_check();
    if (st != KAMI_OK) return st;
    unsigned char* base = nullptr;
    st = static_scratch(&base);
    if (st != KAMI_OK) return st;
    int grid = 0;
    hipError_t e = persistent_grid(precision, b_kn, &grid);
    if (e != hipSuccess) return hip_fail(e);
    build_items(g_static.probs, count, g_static.bucket, variant_of_bucket, ei, eo, b_kn,
                g_static.order, g_static.items);
    hipStream_t s = (hipStream_t)stream;
    // From pageable memory the copy has consumed the staging array when the call returns,
    // so the next call may refill it.
    WorkItem* d_items = reinterpret_cast<WorkItem*>(base + kStaticCounterBytes);
    e = hipMemcpyAsync(d_items, g_static.items, (size_t)n_items * sizeof(WorkItem),
                       hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return hip_fail(e);
    // launch_batched resets the counter with a kernel before the persistent launch
    BatchArgs a{d_items, reinterpret_cast<unsigned int*>(base), n_items, o.alpha, o.beta,
                (int32_t)b_kn, queue_shift(n_items, grid)};
    e = launch_batched(precision, a, grid, s);
    if (e != hipSuccess) return hip_fail(e);
    return KAMI_OK;
}

}  // extern "C"
