// Block-sparse (BSR) SpGEMM C = A * B: symbolic (count, scan, fill) and numeric phases.
//
// Reference: src/sparse/spgemm/2d/block_spgemm_2d_half_mma.cu -- spgemm_symbolic_kernel
// (:58-86, one thread per block row with a dense flag array), thrust::exclusive_scan
// (:806), a blocking copy of nnzC (:808), spgemm_symbolic_col_kernel (:88-116, ascending
// column fill) and the numeric kernel block_spgemm_2d_half_mma (:171) that keeps the whole
// matrix in one CTA.  Here every phase is its own device kernel sized for large matrices:
//   count  -- one workgroup per block row; every (A entry, B entry) pair of the row is
//             one thread's work (four dependent loads per row however long it is); the
//             row's output columns are a bitmap in LDS set with LDS atomics, the count is
//             its population, and with a workspace the bitmap is also stored;
//   scan   -- one workgroup scans the counts in place into rowptrC (exclusive, with
//             rowptrC[mb] = nnzb(C)); chunks of 1024 with a running carry; the total also
//             goes to a pinned host word the host waits for (no copy back);
//   fill   -- the stored bitmap (without a workspace: a rebuilt one) has its set bits
//             written in ascending order, each thread placing the bits of one 32-bit word
//             after the prefix popcount of the words before it;
//   numeric -- one wave per output block (i, j): the row i comes from a 64-way search
//             of rowptrC; the lanes binary-search B's sorted rows k of all blocks (i, k)
//             of A at once for column j; A fragments are loaded straight into
//             registers, B blocks (row-major K x N) are transposed through the wave's LDS
//             into the K-contiguous MFMA layout.
// All integer phases are exact and deterministic.  B may have at most 491520 block
// columns (a 60 KiB bitmap).
#include <algorithm>
#include <chrono>

#include "kami_internal.hpp"

namespace kami {

constexpr int kMaxBitmapWords = 15360;  // 491520 block columns: a 60 KiB bitmap

// Staging of one chunk of a row's A entries (256 per pass): the start of B's row named
// by each entry and the exclusive prefix of those rows' lengths.
struct MarkLds {
    int b0[256];
    int pre[256];
    int wave_tot[4];
};

// Exclusive prefix over the workgroup's 256 threads; the total goes to *total.
__device__ inline int block_exclusive_scan(int v, MarkLds& s, int* total) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int incl = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(incl, off);
        if (lane >= off) incl += y;
    }
    if (lane == 63) s.wave_tot[wave] = incl;
    __syncthreads();
    int before = incl - v;
    for (int q = 0; q < wave; ++q) before += s.wave_tot[q];
    *total = s.wave_tot[0] + s.wave_tot[1] + s.wave_tot[2] + s.wave_tot[3];
    return before;
}

// Mark the output columns of block row `row` in the LDS bitmap (which must be zero).
// Every (A entry, B entry) pair of the row is one unit of work spread over the
// workgroup, so the dependent loads are four in all (A's row range; the entry's column
// and B's row range; the pair's B column; the mark), however many entries the row has:
// thread e takes A entry a0 + e of a chunk of 256, the chunk's B row lengths are scanned,
// and thread p finds the entry of pair p by binary search over the prefix.
__device__ inline void mark_row(int row, const int32_t* rpA, const int32_t* ciA,
                                const int32_t* rpB, const int32_t* ciB, unsigned* bitmap,
                                MarkLds& s) {
    const int tid = threadIdx.x;
    const int a0 = rpA[row], a1 = rpA[row + 1];
    for (int base = a0; base < a1; base += 256) {
        const int n = min(256, a1 - base);
        int b0 = 0, len = 0;
        if (tid < n) {
            const int k = ciA[base + tid];
            b0 = rpB[k];
            len = rpB[k + 1] - b0;
        }
        int pairs;
        const int pre = block_exclusive_scan(len, s, &pairs);
        s.b0[tid] = b0;
        s.pre[tid] = pre;
        __syncthreads();
        for (int p = tid; p < pairs; p += 256) {
            int lo = 0, hi = n - 1;  // the last entry whose prefix is <= p
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (s.pre[mid] <= p) lo = mid; else hi = mid - 1;
            }
            const int j = ciB[s.b0[lo] + p - s.pre[lo]];
            atomicOr(&bitmap[j >> 5], 1u << (j & 31));
        }
        __syncthreads();  // s is rewritten by the next chunk
    }
}

// Per-row output block counts of C = A * B.  With `bitmaps` (mb rows of nwords words,
// the caller's workspace) the row's column bitmap is also stored, for the fill.
__global__ void __launch_bounds__(256)
spgemm_count_kernel(int mb, int nwords, const int32_t* rpA, const int32_t* ciA,
                    const int32_t* rpB, const int32_t* ciB, int32_t* counts, unsigned* bitmaps) {
    extern __shared__ unsigned bitmap[];
    __shared__ MarkLds s;
    __shared__ int s_part[4];
    for (int row = blockIdx.x; row < mb; row += gridDim.x) {
        for (int w = threadIdx.x; w < nwords; w += blockDim.x) bitmap[w] = 0u;
        __syncthreads();
        mark_row(row, rpA, ciA, rpB, ciB, bitmap, s);
        int c = 0;
        for (int w = threadIdx.x; w < nwords; w += blockDim.x) {
            const unsigned bits = bitmap[w];
            c += __popc(bits);
            if (bitmaps) bitmaps[(int64_t)row * nwords + w] = bits;
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) c += __shfl_xor(c, off);
        if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = c;
        __syncthreads();
        if (threadIdx.x == 0) counts[row] = s_part[0] + s_part[1] + s_part[2] + s_part[3];
        __syncthreads();
    }
}

// In-place exclusive scan of x[0..n-2] into x[0..n-1] (x[n-1] receives the total;
// its input value is ignored).  One workgroup of 1024 threads.
// Exclusive scan of the n - 1 per-row counts in x into x[0 .. n-1].  The running total is
// kept in 64 bits and every prefix is stored saturated at INT32_MAX, so an output with
// 2^31 - 1 or more blocks (which int32 row pointers cannot hold) shows as x[n-1] ==
// INT32_MAX instead of wrapping.
// `done`, when given, is a pair of host-mapped words: done[0] receives x[n-1] and then,
// once that store is visible to the host, done[1] receives `seq`; the host waits for seq
// instead of copying the total back.
__global__ void __launch_bounds__(1024) spgemm_scan_kernel(int32_t* x, int n, int32_t* done, int32_t seq) {
    __shared__ int s_wave[16];
    __shared__ long long s_carry;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) s_carry = 0;
    __syncthreads();
    for (int base = 0; base < n; base += 1024) {
        const int i = base + tid;
        const int v = (i < n - 1) ? x[i] : 0;
        int incl = v;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int y = __shfl_up(incl, off);
            if (lane >= off) incl += y;
        }
        if (lane == 63) s_wave[wave] = incl;
        __syncthreads();
        if (wave == 0) {
            int w = lane < 16 ? s_wave[lane] : 0;
#pragma unroll
            for (int off = 1; off < 16; off <<= 1) {
                const int y = __shfl_up(w, off);
                if (lane >= off) w += y;
            }
            if (lane < 16) s_wave[lane] = w;  // inclusive wave totals
        }
        __syncthreads();
        const long long carry = s_carry;
        const int before = (wave > 0 ? s_wave[wave - 1] : 0) + incl - v;  // < 2^30 per chunk
        if (i < n) x[i] = (int32_t)min(carry + before, (long long)INT32_MAX);
        __syncthreads();
        if (tid == 0) s_carry = carry + s_wave[15];
        __syncthreads();
    }
    // the running total is the sum of all counts, x[n - 1] before saturation
    if (done && tid == 0) {
        __hip_atomic_store(&done[0], (int32_t)min(s_carry, (long long)INT32_MAX), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(&done[1], seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

__global__ void __launch_bounds__(256)
spgemm_fill_kernel(int mb, int nwords, const int32_t* rpA, const int32_t* ciA,
                   const int32_t* rpB, const int32_t* ciB, const int32_t* rpC, int32_t* ciC,
                   const unsigned* bitmaps) {
    extern __shared__ unsigned bitmap[];
    __shared__ MarkLds s;
    __shared__ int s_wave[4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int row = blockIdx.x; row < mb; row += gridDim.x) {
        if (bitmaps) {  // the count kernel's bitmap of the row
            for (int w = tid; w < nwords; w += blockDim.x) bitmap[w] = bitmaps[(int64_t)row * nwords + w];
        } else {
            for (int w = tid; w < nwords; w += blockDim.x) bitmap[w] = 0u;
            __syncthreads();
            mark_row(row, rpA, ciA, rpB, ciB, bitmap, s);
        }
        __syncthreads();
        int out = rpC[row];
        for (int base = 0; base < nwords; base += 256) {
            const int w = base + tid;
            const unsigned bits = w < nwords ? bitmap[w] : 0u;
            const int v = __popc(bits);
            int incl = v;
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                const int y = __shfl_up(incl, off);
                if (lane >= off) incl += y;
            }
            if (lane == 63) s_wave[wave] = incl;
            __syncthreads();
            int before = incl - v;
            for (int q = 0; q < wave; ++q) before += s_wave[q];
            int pos = out + before;
            unsigned b = bits;
            while (b) {
                const int bit = __ffs(b) - 1;
                ciC[pos++] = (w << 5) + bit;
                b &= b - 1;
            }
            const int tot = s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
            __syncthreads();
            out += tot;
        }
        __syncthreads();
    }
}

// ---- numeric: one wave per output block ----

template <int BS, class AT>
struct SpNumGeom {
    static constexpr int IM = BS / 16;                 // atom rows (= atom cols)
    static constexpr int KSTEPS = BS / AT::KSTEP;
    static constexpr int ELT = (int)sizeof(typename AT::in_t);
    static constexpr int ROWB = lds_row_bytes(BS * ELT, AT::FRAG_BYTES);  // transposed B row
    static constexpr int LDS = BS * ROWB;
    static constexpr int CHUNKS = BS * BS * ELT / 16;   // 16-byte pieces of a block
    static constexpr int PER_LANE = (CHUNKS + 63) / 64;
};

// The numeric kernel's workgroup is one wave: its LDS accesses complete in order, so
// sB needs no workgroup barrier, only this wave's LDS operations finished and a
// compiler fence.  (__syncthreads() would also wait for every global load in flight.)
__device__ inline void wave_lds_sync() {
    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0); vmcnt and expcnt left at their maxima
    __builtin_amdgcn_wave_barrier();
}

__device__ inline int find_col(const int32_t* ci, int lo, int hi, int j) {
    // index in [lo, hi) with ci[idx] == j, or -1 (ci ascending within the row)
    const int end = hi;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (ci[mid] < j) lo = mid + 1; else hi = mid;
    }
    return (lo < end && ci[lo] == j) ? lo : -1;
}

template <int BS, class AT>
__global__ void __launch_bounds__(64)
spgemm_numeric_kernel(int mb, int64_t nnzc, const int32_t* rpA, const int32_t* ciA,
                      const void* vA_, const int32_t* rpB, const int32_t* ciB, const void* vB_,
                      const int32_t* rpC, const int32_t* ciC, void* vC_, double alpha_d,
                      int iters) {
    using G = SpNumGeom<BS, AT>;
    using in_t = typename AT::in_t;
    using acc_t = typename AT::acc_t;
    using out_t = typename AT::out_t;
    using frag_t = typename AT::frag_t;
    using accv_t = typename AT::accv_t;
    constexpr int FB = AT::FRAG_BYTES;
    __shared__ __attribute__((aligned(16))) unsigned char sB[G::LDS];
    const in_t* vA = reinterpret_cast<const in_t*>(vA_);
    const in_t* vB = reinterpret_cast<const in_t*>(vB_);
    out_t* vC = reinterpret_cast<out_t*>(vC_);
    const int lane = threadIdx.x;
    const acc_t alpha = (acc_t)alpha_d;

    for (int64_t cb = blockIdx.x; cb < nnzc; cb += gridDim.x) {
        const int j = ciC[cb];  // issued before the row search, which does not need it
        // block row of output block cb: the last row with rpC[row] <= cb, found by a
        // 64-way search (each round the lanes probe 64 evenly spaced row pointers)
        int lo = 0, hi = mb;
        while (hi - lo > 1) {
            const int step = (hi - lo + 63) / 64;
            const int idx = lo + (lane + 1) * step;
            const bool le = idx < hi && (int64_t)rpC[idx] <= cb;
            const int c = __popcll(__ballot(le));  // probes are monotone: a prefix is set
            const int nlo = lo + c * step;
            hi = min(hi, nlo + step);
            lo = nlo;
        }
        const int i = lo;
        accv_t acc[G::IM][G::IM];
#pragma unroll
        for (int a = 0; a < G::IM; ++a)
#pragma unroll
            for (int b = 0; b < G::IM; ++b)
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[a][b][v] = 0;

        // every lane searches the B row of one block of A's row i for column j; the
        // matches are then multiplied in ascending order of A's blocks
        const int a0 = rpA[i], a1 = rpA[i + 1];
        for (int base = a0; base < a1; base += 64) {
          const int tl = base + lane;
          int lmatch = -1;
          if (tl < a1) {
              const int k = ciA[tl];
              lmatch = find_col(ciB, rpB[k], rpB[k + 1], j);
          }
          uint64_t todo = __ballot(lmatch >= 0);
          while (todo) {
            const int src = __ffsll((unsigned long long)todo) - 1;
            todo &= todo - 1;
            const int t = base + src;
            const int l = __shfl(lmatch, src);
            const in_t* Ablk = vA + (int64_t)t * BS * BS;
            const in_t* Bblk = vB + (int64_t)l * BS * BS;
            // A fragments straight to registers, issued with the B block's loads
            frag_t af[G::KSTEPS][G::IM], bf[G::KSTEPS][G::IM];
#pragma unroll
            for (int ks = 0; ks < G::KSTEPS; ++ks) {
                const int kk = ks * AT::KSTEP + AT::KV * (lane >> 4);
#pragma unroll
                for (int x = 0; x < G::IM; ++x)
                    af[ks][x] = bitcast<frag_t>(load_raw<FB>(Ablk + (x * 16 + (lane & 15)) * BS + kk));
            }
            // B block (row-major K x N) -> LDS as N rows of K (transposed)
#pragma unroll
            for (int q = 0; q < G::PER_LANE; ++q) {
                const int c = lane + 64 * q;
                if (c >= G::CHUNKS) break;
                constexpr int CPR = BS * G::ELT / 16;   // pieces per block row
                const int kr = c / CPR, jc = (c % CPR) * (16 / G::ELT);
                const u4_t raw = load_raw<16>(Bblk + kr * BS + jc);
                const in_t* e = reinterpret_cast<const in_t*>(&raw);
#pragma unroll
                for (int x = 0; x < 16 / G::ELT; ++x)
                    *reinterpret_cast<in_t*>(sB + (jc + x) * G::ROWB + kr * G::ELT) = e[x];
            }
            wave_lds_sync();
            // both blocks' fragments in registers; `iters` > 1 repeats the products on them
            // (the reference's in-kernel NUM_ITER repeat), iters = 1 is the plain product
#pragma unroll
            for (int ks = 0; ks < G::KSTEPS; ++ks) {
                const int kk = ks * AT::KSTEP + AT::KV * (lane >> 4);
#pragma unroll
                for (int x = 0; x < G::IM; ++x)
                    bf[ks][x] = bitcast<frag_t>(
                        load_raw<FB>(sB + (x * 16 + (lane & 15)) * G::ROWB + kk * G::ELT));
            }
            wave_lds_sync();  // the fragments are read before the next match rewrites sB
            for (int it = 0; it < iters; ++it)
#pragma unroll
                for (int ks = 0; ks < G::KSTEPS; ++ks)
#pragma unroll
                    for (int jn = 0; jn < G::IM; ++jn)
#pragma unroll
                        for (int im = 0; im < G::IM; ++im) AT::mma(acc[im][jn], af[ks][im], bf[ks][jn]);
          }
        }
        out_t* Cblk = vC + cb * BS * BS;
#pragma unroll
        for (int im = 0; im < G::IM; ++im)
#pragma unroll
            for (int jn = 0; jn < G::IM; ++jn)
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    Cblk[(im * 16 + AT::out_row(lane, v)) * BS + jn * 16 + (lane & 15)] =
                        (out_t)(alpha * acc[im][jn][v]);
    }
}

static kami_status check_bsr(const kami_bsr* X) {
    if (!X || X->mb <= 0 || X->nb <= 0 || X->nnzb < 0) return KAMI_EINVAL;
    if (!X->rowptr || (X->nnzb > 0 && (!X->colidx))) return KAMI_EINVAL;
    return KAMI_OK;
}

static kami_status check_pair(const kami_bsr* A, const kami_bsr* B) {
    kami_status st = check_bsr(A);
    if (st == KAMI_OK) st = check_bsr(B);
    if (st != KAMI_OK) return st;
    if (A->nb != B->mb || A->bs != B->bs) return KAMI_EINVAL;
    if (B->nb > kMaxBitmapWords * 32) return KAMI_ENOTSUP;
    return KAMI_OK;
}

// Bytes of the per-row column bitmaps of C = A * B (A with mb block rows, B with nb
// block columns).
static size_t bitmap_bytes(int32_t mb, int32_t nb) {
    return (size_t)mb * (((size_t)nb + 31) / 32) * sizeof(unsigned);
}

// Two pinned, device-visible host words per host thread for the symbolic phase's total
// and its sequence number (the scan kernel writes both; no copy back).  Kept until the
// process ends; null when pinned memory cannot be had.
static int32_t* pinned_total() {
    thread_local int32_t* word = nullptr;
    thread_local bool tried = false;
    if (!tried) {
        tried = true;
        void* p = nullptr;
        if (hipHostMalloc(&p, 64, hipHostMallocPortable | hipHostMallocMapped) == hipSuccess)
            word = static_cast<int32_t*>(p);
        else
            (void)hipGetLastError();
    }
    return word;
}

static int32_t* pinned_device(int32_t* host) {
    void* d = nullptr;
    if (hipHostGetDevicePointer(&d, host, 0) != hipSuccess) {
        (void)hipGetLastError();
        return host;  // unified addressing: the host address is the device address
    }
    return static_cast<int32_t*>(d);
}

// Waits until the device has stored `seq` in the pinned word: a short spin on the word
// (the kernels before it take microseconds), then the stream's own blocking wait, so that
// a long queue of earlier work on the stream does not keep a host core spinning.
static hipError_t wait_for_word(const int32_t* word, int32_t seq, hipStream_t s) {
    const auto t0 = std::chrono::steady_clock::now();
    while (__atomic_load_n(word, __ATOMIC_ACQUIRE) != seq) {
        if (std::chrono::steady_clock::now() - t0 > std::chrono::microseconds(200)) {
            const hipError_t e = hipStreamSynchronize(s);
            if (e != hipSuccess) return e;
            // the stream is done, so the scan kernel has stored the word
            return __atomic_load_n(word, __ATOMIC_ACQUIRE) == seq ? hipSuccess : hipErrorUnknown;
        }
    }
    return hipSuccess;
}

static int sym_grid(int mb) {
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess)
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    return std::max(1, std::min(mb, cus * 8));
}

}  // namespace kami

using namespace kami;

extern "C" {

kami_status kami_bsr_spgemm_workspace_size(int32_t mb, int32_t nb, size_t* bytes) {
    if (mb <= 0 || nb <= 0 || !bytes) return KAMI_EINVAL;
    *bytes = bitmap_bytes(mb, nb);  // one column bitmap per block row of C
    return KAMI_OK;
}

kami_status kami_bsr_spgemm_symbolic(const kami_bsr* A, const kami_bsr* B, int32_t* rowptrC,
                                     int64_t* nnzbC, void* ws, size_t ws_bytes, void* stream) {
    kami_status st = check_pair(A, B);
    if (st != KAMI_OK) return st;
    if (!rowptrC || !nnzbC) return KAMI_EINVAL;
    st = device_check();
    if (st != KAMI_OK) return st;
    hipStream_t s = (hipStream_t)stream;
    const int nwords = (B->nb + 31) / 32;
    unsigned* bitmaps = ws && ws_bytes >= bitmap_bytes(A->mb, B->nb) ? static_cast<unsigned*>(ws) : nullptr;
    int32_t* words = pinned_total();
    thread_local int32_t seq = 0;
    seq = seq == INT32_MAX ? 1 : seq + 1;
    hipLaunchKernelGGL(spgemm_count_kernel, dim3(sym_grid(A->mb)), dim3(256), nwords * 4, s, A->mb,
                       nwords, A->rowptr, A->colidx, B->rowptr, B->colidx, rowptrC, bitmaps);
    hipLaunchKernelGGL(spgemm_scan_kernel, dim3(1), dim3(1024), 0, s, rowptrC, A->mb + 1,
                       words ? pinned_device(words) : nullptr, seq);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e);
    int32_t total = 0;
    if (words) {
        e = wait_for_word(words + 1, seq, s);
        if (e != hipSuccess) return hip_fail(e);
        total = __atomic_load_n(words, __ATOMIC_ACQUIRE);
    } else {  // no pinned words: copy the total back
        e = hipMemcpyAsync(&total, rowptrC + A->mb, sizeof(int32_t), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return hip_fail(e);
    }
    if (total == INT32_MAX) return KAMI_ENOTSUP;  // more blocks than int32 row pointers hold
    *nnzbC = total;
    return KAMI_OK;
}

kami_status kami_bsr_spgemm_fill(const kami_bsr* A, const kami_bsr* B, const int32_t* rowptrC,
                                 int32_t* colidxC, void* ws, size_t ws_bytes, void* stream) {
    kami_status st = check_pair(A, B);
    if (st != KAMI_OK) return st;
    if (!rowptrC || !colidxC) return KAMI_EINVAL;
    st = device_check();
    if (st != KAMI_OK) return st;
    const int nwords = (B->nb + 31) / 32;
    const unsigned* bitmaps =
        ws && ws_bytes >= bitmap_bytes(A->mb, B->nb) ? static_cast<const unsigned*>(ws) : nullptr;
    hipLaunchKernelGGL(spgemm_fill_kernel, dim3(sym_grid(A->mb)), dim3(256), nwords * 4,
                       (hipStream_t)stream, A->mb, nwords, A->rowptr, A->colidx, B->rowptr,
                       B->colidx, rowptrC, colidxC, bitmaps);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e);
    return KAMI_OK;
}

kami_status kami_bsr_spgemm_numeric(const kami_bsr* A, const kami_bsr* B, const int32_t* rowptrC,
                                    const int32_t* colidxC, int64_t nnzbC, void* valC,
                                    int32_t precision, double alpha, void* stream) {
    return kami_bsr_spgemm_numeric_repeat(A, B, rowptrC, colidxC, nnzbC, valC, precision, alpha,
                                          1, stream);
}

kami_status kami_bsr_spgemm_numeric_repeat(const kami_bsr* A, const kami_bsr* B,
                                           const int32_t* rowptrC, const int32_t* colidxC,
                                           int64_t nnzbC, void* valC, int32_t precision,
                                           double alpha, int32_t iters, void* stream) {
    if (iters < 1) return KAMI_EINVAL;
    kami_status st = check_pair(A, B);
    if (st != KAMI_OK) return st;
    if (nnzbC < 0 || !rowptrC) return KAMI_EINVAL;
    if (nnzbC == 0) return KAMI_OK;
    if (!colidxC || !valC || !A->vals || !B->vals) return KAMI_EINVAL;
    if (((uintptr_t)A->vals % 16) || ((uintptr_t)B->vals % 16) || ((uintptr_t)valC % 16))
        return KAMI_EINVAL;
    if (A->bs != 16 && A->bs != 32) return KAMI_ENOTSUP;
    st = device_check();
    if (st != KAMI_OK) return st;
    hipStream_t s = (hipStream_t)stream;
    const unsigned grid = (unsigned)std::min<int64_t>(nnzbC, 1 << 20);
#define KAMI_SPNUM(BS, AT)                                                                      \
    hipLaunchKernelGGL((spgemm_numeric_kernel<BS, AT>), dim3(grid), dim3(64), 0, s, A->mb,       \
                       nnzbC, A->rowptr, A->colidx, A->vals, B->rowptr, B->colidx, B->vals,      \
                       rowptrC, colidxC, valC, alpha, (int)iters)
    if (precision == KAMI_PREC_F16_F32) {
        if (A->bs == 16) KAMI_SPNUM(16, AtomF16x16); else KAMI_SPNUM(32, AtomF16x32);
    } else if (precision == KAMI_PREC_BF16_F32) {
        if (A->bs == 16) KAMI_SPNUM(16, AtomBF16x16); else KAMI_SPNUM(32, AtomBF16x32);
    } else if (precision == KAMI_PREC_F64) {
        if (A->bs == 16) KAMI_SPNUM(16, AtomF64); else KAMI_SPNUM(32, AtomF64);
    } else {
        return KAMI_ENOTSUP;
    }
#undef KAMI_SPNUM
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e);
    return KAMI_OK;
}

}  // extern "C"
