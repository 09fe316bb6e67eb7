// Persistent, deeply prefetched form of the 2D block GEMM.
//
// Same partition as block_gemm_kernel's 2D instances (PM x PN waves, operands shared
// through LDS), but a workgroup walks a sequence of problems (blockIdx.x, + gridDim.x,
// ...) as one stream of K stages.  Each thread keeps PD stages of its staging pieces in
// a register ring, so global loads run PD stages ahead of the MFMAs -- including the
// first stages of the next problem while the current one finishes and stores C.  LDS is
// double-buffered; one barrier per stage.  The stage body is unrolled over the NS stages
// of a problem so ring slots and prefetch targets are compile-time (NS % PD == 0).
#pragma once

#include <algorithm>
#include <atomic>

#include "block_gemm_kernel.hpp"

namespace kami {

template <class AT, class G, int PD>
__global__ void __launch_bounds__(G::NT)
block_gemm_pipe_kernel(BlockArgs args, int64_t count) {
    static_assert(G::PK == 1 && !G::A_PRIV, "pipelined form is the 2D partition");
    static_assert(G::DIM == 16, "pipelined form takes 16 x 16 atoms");
    static_assert(PD >= 1 && G::NS % PD == 0, "prefetch depth must divide the stage count");
    using in_t = typename AT::in_t;
    using acc_t = typename AT::acc_t;
    using out_t = typename AT::out_t;
    using frag_t = typename AT::frag_t;
    using accv_t = typename AT::accv_t;
    constexpr int FB = AT::FRAG_BYTES;
    constexpr int CPR = G::KS / G::VE;  // 16-byte pieces per staged row

    __shared__ __attribute__((aligned(16))) unsigned char smem[2 * G::STAGE_LDS];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int pn = wave % G::PN, pm = wave / G::PN;
    const int64_t lda = args.lda, ldb = args.ldb, ldc = args.ldc;
    const int64_t grid = gridDim.x;
    const in_t* const A0 = reinterpret_cast<const in_t*>(args.A);
    const in_t* const B0 = reinterpret_cast<const in_t*>(args.B);
    out_t* const C0 = reinterpret_cast<out_t*>(args.C);

    u4_t ringA[PD][G::NA], ringB[PD][G::NBC];

    // issue the loads of stage `ks` of problem `p` into ring slot `slot`
    auto load = [&](int64_t p, int ks, u4_t (&ra)[G::NA], u4_t (&rb)[G::NBC]) {
        const in_t* A = A0 + p * args.sa;
        const in_t* B = B0 + p * args.sb;
        const int k0 = ks * G::KS;
#pragma unroll
        for (int i = 0; i < G::NA; ++i) {
            const int c = tid + i * G::NT;
            if (c < G::CA) ra[i] = load_raw<16>(A + (c / CPR) * lda + k0 + (c % CPR) * G::VE);
        }
#pragma unroll
        for (int i = 0; i < G::NBC; ++i) {
            const int c = tid + i * G::NT;
            if (c < G::CB) rb[i] = load_raw<16>(B + (c / CPR) * ldb + k0 + (c % CPR) * G::VE);
        }
    };
    auto write = [&](int buf, const u4_t (&ra)[G::NA], const u4_t (&rb)[G::NBC]) {
        unsigned char* base = smem + buf * G::STAGE_LDS;
#pragma unroll
        for (int i = 0; i < G::NA; ++i) {
            const int c = tid + i * G::NT;
            if (c < G::CA) store_raw<16>(base + (c / CPR) * G::ROWB + (c % CPR) * 16, ra[i]);
        }
#pragma unroll
        for (int i = 0; i < G::NBC; ++i) {
            const int c = tid + i * G::NT;
            if (c < G::CB)
                store_raw<16>(base + G::A_LDS + (c / CPR) * G::ROWB + (c % CPR) * 16, rb[i]);
        }
    };

    accv_t acc[G::IM][G::JN];
    auto compute = [&](int buf) {
        const unsigned char* base = smem + buf * G::STAGE_LDS;
        const unsigned char* bb = base + G::A_LDS;
#pragma unroll
        for (int ks = 0; ks < G::KSTEPS; ++ks) {
            const int kk = (ks * AT::KSTEP + AT::KV * (lane >> 4)) * G::ELT;
            frag_t af[G::IM];
#pragma unroll
            for (int im = 0; im < G::IM; ++im)
                af[im] = bitcast<frag_t>(
                    load_raw<FB>(base + (pm * G::TM + im * 16 + (lane & 15)) * G::ROWB + kk));
            frag_t bf[G::JN];
#pragma unroll
            for (int jn = 0; jn < G::JN; ++jn)
                bf[jn] = bitcast<frag_t>(
                    load_raw<FB>(bb + (pn * G::TN + jn * 16 + (lane & 15)) * G::ROWB + kk));
#pragma unroll
            for (int e = 0; e < AT::NSUB; ++e)
#pragma unroll
                for (int jn = 0; jn < G::JN; ++jn)
#pragma unroll
                    for (int im = 0; im < G::IM; ++im) AT::mma_sub(acc[im][jn], af[im], bf[jn], e);
        }
    };

    const acc_t alpha = (acc_t)args.alpha, beta = (acc_t)args.beta;
    int64_t p = blockIdx.x;
    if (p >= count) return;

    // prologue: stage 0 into LDS buffer 0, stages 1..PD into the ring
    load(p, 0, ringA[0], ringB[0]);
    write(0, ringA[0], ringB[0]);
#pragma unroll
    for (int d = 1; d <= PD; ++d) {
        const int64_t pp = p + grid * (d / G::NS);
        if (pp < count) load(pp, d % G::NS, ringA[d % PD], ringB[d % PD]);
    }
    __syncthreads();

    int buf = 0;
    for (; p < count; p += grid) {
#pragma unroll
        for (int i = 0; i < G::IM; ++i)
#pragma unroll
            for (int j = 0; j < G::JN; ++j)
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[i][j][v] = 0;
#pragma unroll
        for (int s = 0; s < G::NS; ++s) {
            compute(buf);
            // stage s + 1 (maybe stage 0 of the next problem) goes to the other buffer
            const int64_t pn1 = p + grid * ((s + 1) / G::NS);
            if (pn1 < count) {
                write(buf ^ 1, ringA[(s + 1) % PD], ringB[(s + 1) % PD]);
                const int64_t pn2 = p + grid * ((s + 1 + PD) / G::NS);
                if (pn2 < count)
                    load(pn2, (s + 1 + PD) % G::NS, ringA[(s + 1) % PD], ringB[(s + 1) % PD]);
            }
            buf ^= 1;
            if (s + 1 == G::NS) {
                // epilogue of problem p (registers only; the next problem's stages are
                // already in flight)
                out_t* C = C0 + p * args.sc;
#pragma unroll
                for (int im = 0; im < G::IM; ++im)
#pragma unroll
                    for (int jn = 0; jn < G::JN; ++jn) {
                        const int col = pn * G::TN + jn * 16 + (lane & 15);
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            const int row = pm * G::TM + im * 16 + AT::out_row(lane, v);
                            out_t* dst = C + (int64_t)row * ldc + col;
                            acc_t val = alpha * acc[im][jn][v];
                            if (beta != (acc_t)0) val += beta * (acc_t)(*dst);
                            *dst = (out_t)val;
                        }
                    }
            }
            __syncthreads();
        }
    }
}

// Workgroups of `kernel` (`threads` per workgroup) that the current device holds at once:
// CUs x occupancy (at least one per CU).  Cached per device in `cache`: the value depends
// only on the kernel and the device, so concurrent first calls store the same value.
constexpr int kMaxCachedDevices = 64;
template <class F>
hipError_t resident_workgroups(F kernel, int threads,
                               std::atomic<int> (&cache)[kMaxCachedDevices], int* out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    const bool cached = dev >= 0 && dev < kMaxCachedDevices;
    int v = cached ? cache[dev].load(std::memory_order_relaxed) : 0;
    if (v == 0) {
        int n = 0, occ = 0;
        e = hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev);
        if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, threads, 0);
        if (e != hipSuccess) return e;
        v = std::max(1, n) * std::max(1, occ);
        if (cached) cache[dev].store(v, std::memory_order_relaxed);
    }
    *out = v;
    return hipSuccess;
}

// Persistent grid: every CU filled to the kernel's occupancy, never more workgroups than
// problems.
template <class AT, class G, int PD>
hipError_t launch_block_gemm_pipe(const BlockArgs& args, int64_t count, hipStream_t stream) {
    static std::atomic<int> cache[kMaxCachedDevices];
    int slots = 0;
    const hipError_t e =
        resident_workgroups(block_gemm_pipe_kernel<AT, G, PD>, G::NT, cache, &slots);
    if (e != hipSuccess) return e;
    const int64_t grid = std::min<int64_t>(count, (int64_t)slots);
    hipLaunchKernelGGL((block_gemm_pipe_kernel<AT, G, PD>), dim3((unsigned)grid), dim3(G::NT), 0,
                       stream, args, count);
    return hipGetLastError();
}

}  // namespace kami
