// FP64 block GEMM through int8 digits on integer MFMA (split precision, integer form).
//
// C = A * B^T in FP64 from byte digits of 62-bit fixed-point operands (the Ozaki scheme:
// Ozaki, Ogita, Oishi, Rump, Numer. Algorithms 59 (2012); on integer matrix units:
// Uchino, Ozaki, Imamura, "Performance enhancement of the Ozaki Scheme on integer matrix
// multiplication unit", 2024).  No counterpart exists in the reference: it is a
// split-precision FP64 path offered beside native FP64 MFMA, on gfx950's int8 MFMA
// (v_mfma_i32_16x16x64_i8, 5.0 POP/s dense, twice the FP16 rate) with exact int32
// accumulation.
//
// Per problem:
//   1. e_i with |A(i,k)| < 2^e_i over the full row of A (f_j for rows of B, B stored
//      N x K);
//   2. X(i,k) = trunc(|A(i,k)| 2^(62 - e_i)) sign(A(i,k)) is a 63-bit signed integer,
//      written as eight balanced byte digits X = sum_{p=0..7} d_p 2^(8 (7 - p)),
//      d_p in [-128, 127] (p = 0 the most significant);
//   3. level sums T_l = sum_{p + q = l} d_p(A) d_q(B)^T over the whole K on int8 MFMA:
//      |T_l| <= 8 K 2^14 < 2^31, so every level is exact in int32 in any order;
//   4. levels l < L are kept (L = 8: 36 digit products) and combined in FP64 by
//      Horner's rule, C(i,j) = alpha 2^(e_i + f_j - 12) sum_l T_l 2^(-8 l) + beta C(i,j).
// Truncating each operand costs at most 2^(e - 62) per element; the dropped levels
// l >= L weigh at most 2^(-8 L) each against the kept ones.  Finite inputs are required.
//
// Work layout (the reference's 1D partition: A private per wave, B broadcast, F64-1D:73-
// 102): a workgroup of BM / 16 waves takes one problem at a time (persistent over
// problems).  Wave w slices rows 16w .. 16w + 15 of A into its own registers -- exactly
// the MFMA A fragments it needs -- and the same rows of B into LDS, where every wave reads
// B fragments for its 16 x BN strip of C.  Lane l of a wave handles operand row
// 16 w + (l mod 16) and, in each 64-wide K step h, the 16 k values
// 64 h + 8 i + 2 (l / 16) + q (i = 0..7, q = 0..1): four lanes read 64 contiguous bytes
// of a row, and slot j = 2 i + q of the lane's 16-byte fragment holds k(i, q).  A and B
// use the same (lane group, slot) -> k map, so each MFMA pairs A(r, k) with B(c, k)
// whatever order the instruction sums its 64 products in.
#include <atomic>
#include <cstdint>

#include "split_common.hpp"

namespace kami {

typedef int i4_t __attribute__((ext_vector_type(4)));

template <int BM_, int BK_, int L_>
struct I8Geom {
    static constexpr int BM = BM_, BN = BM_, BK = BK_;
    static constexpr int S = 8;   // byte digits of a 62-bit fixed-point value
    static constexpr int L = L_;  // levels p + q < L are kept
    static constexpr int WAVES = BM / 16, NT = WAVES * kWave;
    static constexpr int JN = BN / 16, KS = BK / 64;
    static constexpr int FRAG = kWave * 16;  // one wave's 16-byte pieces
    // Region jn of LDS holds the digit planes of B rows 16 jn .. 16 jn + 15 (plane p,
    // K step h at (p KS + h) FRAG) and, before they are sliced, the same rows' raw
    // doubles (K step h, piece i at (8 h + i) FRAG): both take S KS FRAG bytes.
    static constexpr int REGION = S * KS * FRAG;
    static constexpr int EXP_OFF = JN * REGION;
    static constexpr int LDS = EXP_OFF + (BM + BN) * 4;
    static_assert(BM % 16 == 0 && BK % 64 == 0 && WAVES >= 1 && WAVES <= 16, "tiling");
    static_assert(S == 8, "a region holds a K step's eight raw pieces or eight planes");
    static_assert(LDS <= 160 * 1024, "LDS budget");
    static_assert(L >= 1 && L <= 2 * S - 1, "levels");
    // at most S products of two digits (|d| <= 128) per k and level
    static_assert((double)S * BK * 16384.0 < 2147483648.0, "level sums exact in int32");
    static constexpr int products() {
        int n = 0;
        for (int p = 0; p < S; ++p)
            for (int q = 0; q < S; ++q) n += (p + q < L);
        return n;
    }
};

// The eight balanced byte digits of X = trunc(|x| 2^(62 - E)) sign(x), for |x| < 2^E,
// packed in one word: byte b is the digit of weight 2^(8b), in [-128, 127], and
// X = sum_b d_b 2^(8b) exactly.  x is first scaled by 2^-E (exact; it also brings the
// subnormal entries of a tiny row into the normal range); the 53-bit significand of
// x 2^-E with its leading bit placed at bit 62 is then shifted right by
// 1023 - exponent(x 2^-E) >= 1 (a shift of 63 or more gives 0, which covers zero and
// anything below 2^-62 of the row).  Adding 0x80 to every byte with carries and
// flipping each byte's top bit back turns the two's complement value into balanced
// digits.
__device__ inline uint64_t digit_word(double x, int neg_e) {
    const uint64_t bits = __builtin_bit_cast(uint64_t, __builtin_amdgcn_ldexp(x, neg_e));
    const uint32_t h = (uint32_t)(bits >> 32), l = (uint32_t)bits;
    const uint32_t ex = (h >> 20) & 0x7ffu;
    const uint32_t r = min(1023u - ex, 63u);
    const uint32_t mh = (__builtin_amdgcn_alignbit(h, l, 22) & 0x3fffffffu) | 0x40000000u;
    const uint64_t X = ((uint64_t)mh << 32 | (uint64_t)(l << 10)) >> r;
    const uint64_t s = (uint64_t)(int64_t)((int32_t)h >> 31);  // 0 or all ones
    const uint64_t Y = (X ^ s) + (0x8080808080808080ull + (s & 1));  // +-X + 0x80.. per byte
    return Y ^ 0x8080808080808080ull;
}

// Byte transpose of four digit words: out[b] = byte b of w0, w1, w2, w3 in its bytes
// 0..3.  v_perm_b32 (S0, S1, sel) picks bytes of {S0:S1}: 0-3 from S1, 4-7 from S0.
__device__ inline void transpose4(uint64_t w0, uint64_t w1, uint64_t w2, uint64_t w3,
                                  uint32_t (&out)[8]) {
#pragma unroll
    for (int half = 0; half < 2; ++half) {
        const uint32_t a = (uint32_t)(w0 >> (32 * half)), b = (uint32_t)(w1 >> (32 * half));
        const uint32_t c = (uint32_t)(w2 >> (32 * half)), d = (uint32_t)(w3 >> (32 * half));
        const uint32_t ab_lo = __builtin_amdgcn_perm(b, a, 0x05010400u);  // a0 b0 a1 b1
        const uint32_t ab_hi = __builtin_amdgcn_perm(b, a, 0x07030602u);  // a2 b2 a3 b3
        const uint32_t cd_lo = __builtin_amdgcn_perm(d, c, 0x05010400u);
        const uint32_t cd_hi = __builtin_amdgcn_perm(d, c, 0x07030602u);
        out[4 * half + 0] = __builtin_amdgcn_perm(cd_lo, ab_lo, 0x05040100u);  // a0 b0 c0 d0
        out[4 * half + 1] = __builtin_amdgcn_perm(cd_lo, ab_lo, 0x07060302u);  // a1 b1 c1 d1
        out[4 * half + 2] = __builtin_amdgcn_perm(cd_hi, ab_hi, 0x05040100u);
        out[4 * half + 3] = __builtin_amdgcn_perm(cd_hi, ab_hi, 0x07060302u);
    }
}

// Digit fragments of one 64-wide K step from the lane's eight 16-byte pieces x[i]
// (doubles k(i, 0), k(i, 1)): frag[p] is the 16-byte MFMA fragment of plane p (p = 0 the
// most significant digit), slot j = 2 i + q.
__device__ inline void slice_step(const u4_t (&x)[8], int neg_e, u4_t (&frag)[8]) {
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const d2_t x0 = bitcast<d2_t>(x[2 * m]), x1 = bitcast<d2_t>(x[2 * m + 1]);
        uint32_t out[8];
        transpose4(digit_word(x0[0], neg_e), digit_word(x0[1], neg_e), digit_word(x1[0], neg_e),
                   digit_word(x1[1], neg_e), out);
#pragma unroll
        for (int b = 0; b < 8; ++b) frag[7 - b][m] = out[b];
    }
}

// Exponent e with |x| < 2^e over a row: the largest upper word of |x| over the lane's
// values and the row's four lane groups (lanes l, l ^ 16, l ^ 32, l ^ 48), completed
// with all-ones lower bits so that frexp also covers subnormal rows (an all-zero row
// gives the exponent of the smallest subnormal range, and digits 0).
template <int KS>
__device__ inline int row_exponent(const u4_t (&x)[KS][8]) {
    uint32_t m = 0;
#pragma unroll
    for (int h = 0; h < KS; ++h)
#pragma unroll
        for (int i = 0; i < 8; ++i) m = max(m, max(x[h][i][1] & 0x7fffffffu, x[h][i][3] & 0x7fffffffu));
    m = max(m, (uint32_t)__shfl_xor((int)m, 16, 64));
    m = max(m, (uint32_t)__shfl_xor((int)m, 32, 64));
    return __builtin_amdgcn_frexp_exp(__builtin_bit_cast(double, ((uint64_t)m << 32) | 0xffffffffu));
}

// Persistent over problems p = blockIdx.x, + gridDim.x, ...  REPEAT (iters > 1) repeats
// the products on the sliced digits (the reference's resident-repeat NUM_ITER loop) and
// accumulates C = alpha iters A B^T + beta C.  Operand and result addresses are a
// problem's base plus 32-bit byte offsets (the host checks that a problem's rows span
// less than 2^31 bytes).
//
// Pipeline.  The raw rows of A, of B and the wave's A digits do not fit in registers
// together beside the level sums, so the next problem's B rows go straight to LDS: once
// every wave has finished column block jn of the current products (a barrier per block),
// wave jn loads B rows 16 jn .. 16 jn + 15 of the next problem into region jn with
// direct-to-LDS loads.  The next problem's A rows are loaded into registers right after
// the current ones are sliced.  Slicing a problem: A from registers into the wave's
// fragments, then the wave's raw B rows from its own region into registers and their
// digit planes back into the same region.
template <class G, bool REPEAT>
__global__ void __launch_bounds__(G::NT)
split_i8_kernel(SplitArgs args, int64_t count) {
    __shared__ __attribute__((aligned(16))) unsigned char smem[G::LDS];
    int* const expo = reinterpret_cast<int*>(smem + G::EXP_OFF);  // [BM] of A, then [BN] of B

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r = lane & 15, grp = lane >> 4;
    const int row = wave * 16 + r;  // operand row this lane slices, of A and of B
    const int64_t grid = gridDim.x;
    const uint32_t a_off = (uint32_t)((row * args.lda + 2 * grp) * 8);
    const uint32_t b_off = (uint32_t)((row * args.ldb + 2 * grp) * 8);
    // byte offset of the lane's output (row 16 wave + 4 grp + v, column 16 jn + r)
    const uint32_t c_off = (uint32_t)(((wave * 16 + 4 * grp) * args.ldc + r) * 8);
    const uint32_t c_row = (uint32_t)(args.ldc * 8);
    unsigned char* const region = smem + wave * G::REGION;  // this wave's B rows

    u4_t xa[G::KS][8];
    // A rows of problem q through a buffer descriptor; with `live` false the descriptor
    // has no range and the loads return zeros without touching memory, so the loads are
    // unconditional (no branch around the registers they fill)
    auto load_a = [&](int64_t q, bool live) {
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<double*>(args.A + q * args.sa), (short)0, live ? 0x7FFFFFF0 : 0, 0x00020000);
#pragma unroll
        for (int h = 0; h < G::KS; ++h)
#pragma unroll
            for (int i = 0; i < 8; ++i)
                xa[h][i] = __builtin_bit_cast(u4_t, __builtin_amdgcn_raw_buffer_load_b128(
                                                        rs, (int)(a_off + 8 * (64 * h + 8 * i)), 0, 0));
    };
    // the wave's B rows of problem q into its region, piece (h, i) of lane l at
    // (8 h + i) FRAG + 16 l
    auto load_b_lds = [&](int64_t q) {
        const unsigned char* src = reinterpret_cast<const unsigned char*>(args.B + q * args.sb);
#pragma unroll
        for (int h = 0; h < G::KS; ++h)
#pragma unroll
            for (int i = 0; i < 8; ++i)
                __builtin_amdgcn_global_load_lds(
                    src + b_off + 8 * (64 * h + 8 * i),
                    (__attribute__((address_space(3))) void*)(region + (8 * h + i) * G::FRAG), 16, 0,
                    0);
    };
    // B fragment of plane p, column block jn, K step h
    auto bfrag_at = [&](int p, int jn, int h) {
        return smem + jn * G::REGION + (p * G::KS + h) * G::FRAG + 16 * lane;
    };

    int64_t p = blockIdx.x;
    if (p >= count) return;  // uniform
    load_b_lds(p);
    load_a(p, true);
    u4_t ad[G::S][G::KS];  // the wave's A digit fragments

    for (; p < count; p += grid) {
        const bool next = p + grid < count;
        // ---- A: exponents and digits into registers ----
        const int ea = row_exponent<G::KS>(xa);
        expo[row] = ea;  // the row's four lane groups write the same value
#pragma unroll
        for (int h = 0; h < G::KS; ++h) {
            u4_t f[8];
            slice_step(xa[h], -ea, f);
#pragma unroll
            for (int q = 0; q < G::S; ++q) ad[q][h] = f[q];
            __builtin_amdgcn_sched_barrier(0);
        }
        // ---- B: the wave's raw rows from its region, digit planes back into it ----
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): the region's loads have landed
        {
            u4_t xb[G::KS][8];
#pragma unroll
            for (int h = 0; h < G::KS; ++h)
#pragma unroll
                for (int i = 0; i < 8; ++i) xb[h][i] = load_raw<16>(region + (8 * h + i) * G::FRAG + 16 * lane);
            const int eb = row_exponent<G::KS>(xb);
            expo[G::BM + row] = eb;
#pragma unroll
            for (int h = 0; h < G::KS; ++h) {
                u4_t f[8];
                slice_step(xb[h], -eb, f);
#pragma unroll
                for (int q = 0; q < G::S; ++q) store_raw<16>(bfrag_at(q, wave, h), f[q]);
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        // the next problem's A rows are in flight during the products
        load_a(next ? p + grid : p, next);
        __syncthreads();

        // ---- level sums on int8 MFMA, Horner in FP64, store ----
        unsigned char* const Cp = reinterpret_cast<unsigned char*>(args.C + p * args.sc);
        int erow[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) erow[v] = expo[wave * 16 + 4 * grp + v];
#pragma unroll 1
        for (int jn = 0; jn < G::JN; ++jn) {
            const int ecol = expo[G::BM + jn * 16 + r];
            double sum[4] = {0.0, 0.0, 0.0, 0.0};
            for (int it = 0; it < (REPEAT ? args.iters : 1); ++it) {
                i4_t T[G::L];
#pragma unroll
                for (int l = 0; l < G::L; ++l) T[l] = i4_t{0, 0, 0, 0};
                // B fragments one ahead of the products that use them
                i4_t bf = bitcast<i4_t>(load_raw<16>(bfrag_at(0, jn, 0)));
#pragma unroll
                for (int h = 0; h < G::KS; ++h)
#pragma unroll
                    for (int pb = 0; pb < G::L && pb < G::S; ++pb) {
                        const bool last = h + 1 == G::KS && (pb + 1 == G::L || pb + 1 == G::S);
                        i4_t bn = bf;
                        if (!last)
                            bn = bitcast<i4_t>(load_raw<16>(
                                pb + 1 < G::L && pb + 1 < G::S ? bfrag_at(pb + 1, jn, h)
                                                                : bfrag_at(0, jn, h + 1)));
#pragma unroll
                        for (int pa = 0; pa + pb < G::L && pa < G::S; ++pa)
                            T[pa + pb] = __builtin_amdgcn_mfma_i32_16x16x64_i8(
                                bitcast<i4_t>(ad[pa][h]), bf, T[pa + pb], 0, 0, 0);
                        bf = bn;
                        __builtin_amdgcn_sched_barrier(0);
                    }
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    double acc = (double)T[G::L - 1][v];
#pragma unroll
                    for (int l = G::L - 2; l >= 0; --l)
                        acc = __builtin_fma(acc, 1.0 / 256.0, (double)T[l][v]);
                    sum[v] += acc;
                }
            }
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                double* dst = reinterpret_cast<double*>(Cp + c_off + v * c_row + jn * 128);
                double val = __builtin_amdgcn_ldexp(sum[v], erow[v] + ecol - 12) * args.alpha;
                if (args.beta != 0.0) val += args.beta * (*dst);
                *dst = val;
            }
            // every wave is done with region jn: its wave loads the next problem's rows
            __syncthreads();
            if (next && wave == jn) load_b_lds(p + grid);
        }
    }
}

template <class G, bool REPEAT>
hipError_t launch_split_i8_form(const SplitArgs& a, int64_t count, hipStream_t s) {
    static std::atomic<int> cache[kMaxCachedDevices];
    int slots = 0;
    const hipError_t e = resident_workgroups(split_i8_kernel<G, REPEAT>, G::NT, cache, &slots);
    if (e != hipSuccess) return e;
    const int64_t grid = std::min<int64_t>(count, (int64_t)slots);
    hipLaunchKernelGGL((split_i8_kernel<G, REPEAT>), dim3((unsigned)grid), dim3(G::NT), 0, s, a,
                       count);
    return hipGetLastError();
}

template <class G>
hipError_t launch_split_i8(const SplitArgs& a, int64_t count, hipStream_t s) {
    return a.iters > 1 ? launch_split_i8_form<G, true>(a, count, s)
                       : launch_split_i8_form<G, false>(a, count, s);
}

template <int BM, int BK, int L>
static void add_i8(std::vector<SplitInstance>& v) {
    using Geo = I8Geom<BM, BK, L>;
    // persistent (pipe 2) for any iters: the repetitions reuse the sliced digits
    v.push_back(SplitInstance{KAMI_PREC_F64_SPLIT_I8, BM, BM, BK, Geo::S, Geo::WAVES, 1, Geo::NT,
                              Geo::LDS, &launch_split_i8<Geo>, 2, 2});
}

void add_split_i8_instances(std::vector<SplitInstance>& v) {
    add_i8<128, 128, 8>(v);  // 36 digit products, 8 waves, 129 KB of B digit planes
    add_i8<64, 64, 8>(v);    // 4 waves, 33 KB
}

}  // namespace kami
