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
#include <type_traits>
#include <utility>

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
    // A region holds the digit planes of 16 rows of B (plane p, K step h at
    // (p KS + h) FRAG) or, before they are sliced, the same rows' raw doubles (K step h,
    // piece i at (8 h + i) FRAG): both take S KS FRAG bytes.  JN + 1 region slots: the
    // spare slot lets the next problem's first rows land while every slot of the current
    // problem is still being read.
    static constexpr int REGION = S * KS * FRAG;
    static constexpr int SLOTS = JN + 1;
    static constexpr int EXP_OFF = SLOTS * REGION;
    static constexpr int EXP_BUF = (BM + BN) * 4;  // two buffers: problem parity
    static constexpr int LDS = EXP_OFF + 2 * EXP_BUF;
    static_assert(BM % 16 == 0 && BK % 64 == 0 && WAVES >= 2 && WAVES <= 16, "tiling");
    static_assert(S == 8, "a region holds a K step's eight raw pieces or eight planes");
    // the raw pieces of a region (8 per K step) are loaded by all waves, PPW each
    static constexpr int PPW = 8 * KS / WAVES;
    static_assert(PPW >= 1 && PPW * WAVES == 8 * KS, "pieces per wave");
    static_assert(LDS <= 160 * 1024, "LDS budget");
    static_assert(L >= 1 && L <= 2 * S - 1, "levels");
    // at most S products of two digits (|d| <= 128) per k and level
    static_assert((double)S * BK * 16384.0 < 2147483648.0, "level sums exact in int32");
    // repetitions of the products whose summed level sums stay exact in int32
    static constexpr int ACC_ITERS = (int)(2147483647LL / ((long long)S * BK * 16384));
    static_assert(ACC_ITERS >= 1, "one repetition fits");
    static constexpr int products() {
        int n = 0;
        for (int p = 0; p < S; ++p)
            for (int q = 0; q < S; ++q) n += (p + q < L);
        return n;
    }
};

// The eight balanced byte digits of X = trunc(|x| 2^(62 - E)) sign(x), for |x| < 2^E,
// packed in one word: byte b is the digit of weight 2^(8b), in [-128, 127], and
// X = sum_b d_b 2^(8b) exactly.  With c_e = 1023 + E: the 53-bit significand of x with its
// leading bit placed at bit 62 is shifted right by min(c_e - exponent field, 63) (63
// gives 0: bit 63 is clear).  A zero or subnormal x (exponent field 0) is then 0, which
// is exact when E >= -960 (|x| < 2^-1022 <= 2^(E - 62)); rows with smaller E are scaled
// by 2^1000 first (slice_step).  Adding 0x80 to every byte with carries and flipping each
// byte's top bit back turns the two's complement value into balanced digits.
__device__ inline uint64_t digit_word(double x, uint32_t c_e) {
    const uint64_t bits = __builtin_bit_cast(uint64_t, x);
    const uint32_t h = (uint32_t)(bits >> 32), l = (uint32_t)bits;
    const uint32_t r = min(c_e - ((h >> 20) & 0x7ffu), 63u);
    const uint64_t M = ((uint64_t)((h & 0xfffffu) | 0x100000u) << 32 | l) << 10;
    const uint64_t X = M >> r;
    const int32_t s = (int32_t)h >> 31;  // 0 or -1
    const uint64_t s64 = (uint64_t)(int64_t)s;
    // -X = ~X + 1: the +1 rides in the low word of the byte-wise 0x80 offset
    const uint64_t Y = (X ^ s64) + ((uint64_t)0x80808080u << 32 | (uint32_t)(0x80808080u - s));
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

// s_waitcnt immediate that waits for at most n outstanding vector memory operations and
// leaves the other counters alone (gfx9 encoding: vmcnt bits 3:0 and 15:14, expcnt 6:4,
// lgkmcnt 11:8).
constexpr int vmcnt_only(int n) { return ((n >> 4) & 3) << 14 | 0x0F70 | (n & 15); }

// Workgroup barrier for LDS traffic only: this wave's LDS operations complete
// (lgkmcnt(0)), then s_barrier.  __syncthreads() also fences global memory, which waits
// for every store and direct-to-LDS load in flight; the data those loads bring is read
// only by the wave that issued them, after its own vmcnt wait.
__device__ inline void lds_barrier() {
    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0); vmcnt, expcnt left at their maxima
    __builtin_amdgcn_s_barrier();
}

// Rows whose exponent is below this take the scaled path of slice_step.
constexpr int kTinyRowExp = -960;

// Digit fragments of one 64-wide K step from the lane's eight 16-byte pieces x[i]
// (doubles k(i, 0), k(i, 1)): frag[p] is the 16-byte MFMA fragment of plane p (p = 0 the
// most significant digit), slot j = 2 i + q.  `tiny`: some row of the wave has
// E < -960; its elements are scaled by 2^1000 (exact) so that their subnormals are
// normal, with E raised by the same amount.
__device__ inline void slice_step(const u4_t (&x)[8], int E, bool tiny, u4_t (&frag)[8]) {
    const int sc = (tiny && E < kTinyRowExp) ? 1000 : 0;
    const uint32_t c_e = (uint32_t)(1023 + E + sc);
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        d2_t x0 = bitcast<d2_t>(x[2 * m]), x1 = bitcast<d2_t>(x[2 * m + 1]);
        if (tiny) {
            x0[0] = __builtin_amdgcn_ldexp(x0[0], sc);
            x0[1] = __builtin_amdgcn_ldexp(x0[1], sc);
            x1[0] = __builtin_amdgcn_ldexp(x1[0], sc);
            x1[1] = __builtin_amdgcn_ldexp(x1[1], sc);
        }
        uint32_t out[8];
        transpose4(digit_word(x0[0], c_e), digit_word(x0[1], c_e), digit_word(x1[0], c_e),
                   digit_word(x1[1], c_e), out);
#pragma unroll
        for (int b = 0; b < 8; ++b) frag[7 - b][m] = out[b];
    }
}

// Exponent E with |x| < 2^E over a row: frexp of the largest |x| over the lane's values
// and the row's four lane groups (lanes l, l ^ 16, l ^ 32, l ^ 48); 0 for a zero row.
template <int KS>
__device__ inline int row_exponent(const u4_t (&x)[KS][8]) {
    double m = 0.0;
#pragma unroll
    for (int h = 0; h < KS; ++h)
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const d2_t d = bitcast<d2_t>(x[h][i]);
            m = __builtin_fmax(m, __builtin_fmax(__builtin_fabs(d[0]), __builtin_fabs(d[1])));
        }
    m = __builtin_fmax(m, __shfl_xor(m, 16, 64));
    m = __builtin_fmax(m, __shfl_xor(m, 32, 64));
    return __builtin_amdgcn_frexp_exp(m);
}

// Persistent over problems p = blockIdx.x, + gridDim.x, ...  REPEAT (iters > 1) repeats
// the products on the sliced digits (the reference's resident-repeat NUM_ITER loop) and
// accumulates C = alpha iters A B^T + beta C.  Operand and result addresses are a
// problem's base plus 32-bit byte offsets (the host checks that a problem's rows span
// less than 2^31 bytes).
//
// Pipeline.  The raw rows of A, of B and the wave's A digits do not fit in registers
// together beside the level sums, so the next problem's B rows go straight to LDS with
// direct-to-LDS loads, 16 rows (one region) at a time, every wave loading a share of each
// region.  With i the workgroup's problem count, region jn of problem i sits in slot
// (jn - i) mod (JN + 1): region 0 of the next problem takes the slot the current one
// leaves free (loaded as the products start), and region jn of the next problem takes
// the slot of the current one's region jn - 1 once every wave has finished that column
// block (a barrier per block).  The next problem's A rows are loaded into registers
// after the current problem is sliced.  Slicing a problem: A from registers into the
// wave's fragments; then, once every wave's loads have landed (a barrier), the wave's
// raw B rows from region `wave` into registers and their digit planes back into the same
// region.  Barriers wait for LDS operations only (lds_barrier), not for the loads in
// flight.  Exponents alternate between two buffers, so the last column block needs no
// barrier.
template <class G, bool REPEAT>
__global__ void __launch_bounds__(G::NT)
split_i8_kernel(SplitArgs args, int64_t count) {
    __shared__ __attribute__((aligned(16))) unsigned char smem[G::LDS];

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform
    const int r = lane & 15, grp = lane >> 4;
    const int row = wave * 16 + r;  // operand row this lane slices, of A and of B
    const int64_t grid = gridDim.x;
    const uint32_t a_off = (uint32_t)((row * args.lda + 2 * grp) * 8);
    // byte offset of the lane's output (row 16 wave + 4 grp + v, column 16 jn + r)
    const uint32_t c_off = (uint32_t)(((wave * 16 + 4 * grp) * args.ldc + r) * 8);
    const uint32_t c_row = (uint32_t)(args.ldc * 8);

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
    auto slot_of = [&](int64_t i, int jn) {  // region jn of the workgroup's i-th problem
        return smem + (int)(((int64_t)jn - i % G::SLOTS + G::SLOTS) % G::SLOTS) * G::REGION;
    };
    // Raw B rows 16 reg .. 16 reg + 15 of problem q into LDS region `dst` with
    // direct-to-LDS buffer loads: piece pc = 8 h + i (K step h, lane l: doubles
    // 64 h + 8 i + 2 (l / 16) + {0, 1} of row 16 reg + l mod 16) lands at pc FRAG + 16 l.
    // This wave issues pieces [first, first + n).  `live` false: a rangeless descriptor
    // (no memory traffic), so every wave issues the same loads on every path.
    const uint32_t b_lane = (uint32_t)((r * args.ldb + 2 * grp) * 8);
    const uint32_t b_rows16 = (uint32_t)(16 * args.ldb * 8);
    auto load_b_piece = [&](__amdgpu_buffer_rsrc_t rs, uint32_t off, unsigned char* dst, int pc) {
        __builtin_amdgcn_raw_ptr_buffer_load_lds(
            rs, (__attribute__((address_space(3))) void*)(dst + pc * G::FRAG), 16,
            (int)(off + 8 * (64 * (pc >> 3) + 8 * (pc & 7))), 0, 0, 0);
    };
    auto b_rsrc = [&](int64_t q, bool live) {
        return __builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(args.B + q * args.sb), (short)0,
                                                 live ? 0x7FFFFFF0 : 0, 0x00020000);
    };
    // this wave's share of region `reg` of problem q: pieces wave PPW .. wave PPW + PPW - 1
    auto load_b_share = [&](int64_t q, bool live, int reg, unsigned char* dst) {
        const __amdgpu_buffer_rsrc_t rs = b_rsrc(q, live);
        const uint32_t off = b_lane + (uint32_t)reg * b_rows16;
#pragma unroll
        for (int t = 0; t < G::PPW; ++t) load_b_piece(rs, off, dst, wave * G::PPW + t);
    };

    int64_t p = blockIdx.x;
    if (p >= count) return;  // uniform
    {  // the first problem: each wave loads its own rows
        const __amdgpu_buffer_rsrc_t rs = b_rsrc(p, true);
#pragma unroll
        for (int pc = 0; pc < 8 * G::KS; ++pc)
            load_b_piece(rs, b_lane + (uint32_t)wave * b_rows16, slot_of(0, wave), pc);
    }
    load_a(p, true);
    u4_t ad[G::S][G::KS];  // the wave's A digit fragments

    for (int64_t i = 0; p < count; p += grid, ++i) {
        const bool next = p + grid < count;
        int* const expo = reinterpret_cast<int*>(smem + G::EXP_OFF + (i & 1) * G::EXP_BUF);
        unsigned char* const region = slot_of(i, wave);  // this wave's B rows
        // ---- A: exponents and digits into registers ----
        {
            const int ea = row_exponent<G::KS>(xa);
            expo[row] = ea;  // the row's four lane groups write the same value
            const bool tiny = __builtin_amdgcn_ballot_w64(ea < kTinyRowExp) != 0;
#pragma unroll
            for (int h = 0; h < G::KS; ++h) {
                u4_t f[8];
                slice_step(xa[h], ea, tiny, f);
#pragma unroll
                for (int q = 0; q < G::S; ++q) ad[q][h] = f[q];
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        // ---- B: the wave's raw rows from its region, digit planes back into it ----
        // every wave's direct-to-LDS loads have landed
        __builtin_amdgcn_s_waitcnt(vmcnt_only(0));
        lds_barrier();
        {
            u4_t xb[G::KS][8];
#pragma unroll
            for (int h = 0; h < G::KS; ++h)
#pragma unroll
                for (int q = 0; q < 8; ++q) xb[h][q] = load_raw<16>(region + (8 * h + q) * G::FRAG + 16 * lane);
            const int eb = row_exponent<G::KS>(xb);
            expo[G::BM + row] = eb;
            const bool tiny = __builtin_amdgcn_ballot_w64(eb < kTinyRowExp) != 0;
#pragma unroll
            for (int h = 0; h < G::KS; ++h) {
                u4_t f[8];
                slice_step(xb[h], eb, tiny, f);
#pragma unroll
                for (int q = 0; q < G::S; ++q)
                    store_raw<16>(region + (q * G::KS + h) * G::FRAG + 16 * lane, f[q]);
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        // the next problem's A rows: in flight during the products
        load_a(next ? p + grid : p, next);
        lds_barrier();
        // region 0 of the next problem goes to the slot this problem leaves free
        load_b_share(next ? p + grid : p, next, 0, slot_of(i + 1, 0));

        // ---- level sums on int8 MFMA, Horner in FP64, store ----
        unsigned char* const Cp = reinterpret_cast<unsigned char*>(args.C + p * args.sc);
        int erow[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) erow[v] = expo[wave * 16 + 4 * grp + v];
#pragma unroll
        for (int jn = 0; jn < G::JN; ++jn) {
            const int ecol = expo[G::BM + jn * 16 + r];
            const unsigned char* const bsrc = slot_of(i, jn) + 16 * lane;
            auto bfrag = [&](int pb, int h) {
                return bitcast<i4_t>(load_raw<16>(bsrc + (pb * G::KS + h) * G::FRAG));
            };
            // The repetitions accumulate the level sums in int32 (exact while a run of them
            // stays below 2^31: ACC_ITERS repetitions) and are combined in FP64 once per
            // run.  B fragments go in the plane order 0, 7, 1, 6, 2, 5, 3, 4 (8, 1, 7, ...
            // products each), read two ahead through a ring of four; the last two reads of
            // a repetition fetch the next one's first fragments (the same data).
            static_assert(G::L == G::S && G::S == 8, "eight planes, all levels below 8");
            constexpr int kOrder[8] = {0, 7, 1, 6, 2, 5, 3, 4};
            constexpr int NF = 8 * G::KS;
            double sum[4] = {0.0, 0.0, 0.0, 0.0};
            i4_t fb[4];
            fb[0] = bfrag(kOrder[0], 0);
            fb[1] = bfrag(kOrder[1], 0);
            const int iters = REPEAT ? args.iters : 1;
            for (int it = 0; it < iters;) {
                const int run = REPEAT ? min(iters - it, G::ACC_ITERS) : 1;
                i4_t T[G::L];
#pragma unroll
                for (int l = 0; l < G::L; ++l) T[l] = i4_t{0, 0, 0, 0};
                for (int rr = 0; rr < run; ++rr) {
#pragma unroll
                    for (int f = 0; f < NF; ++f) {
                        const int h = f / 8, pb = kOrder[f % 8];
                        fb[(f + 2) % 4] = bfrag(kOrder[(f + 2) % 8], ((f + 2) % NF) / 8);
#pragma unroll
                        for (int pa = 0; pa + pb < G::L; ++pa)
                            T[pa + pb] = __builtin_amdgcn_mfma_i32_16x16x64_i8(
                                bitcast<i4_t>(ad[pa][h]), fb[f % 4], T[pa + pb], 0, 0, 0);
                        __builtin_amdgcn_sched_barrier(0);
                    }
                }
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    double acc = (double)T[G::L - 1][v];
#pragma unroll
                    for (int l = G::L - 2; l >= 0; --l)
                        acc = __builtin_fma(acc, 1.0 / 256.0, (double)T[l][v]);
                    sum[v] += acc;
                    __builtin_amdgcn_sched_barrier(0);  // one output's conversions at a time
                }
                it += run;
            }
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                double* dst = reinterpret_cast<double*>(Cp + c_off + v * c_row + jn * 128);
                double val = __builtin_amdgcn_ldexp(sum[v], erow[v] + ecol - 12) * args.alpha;
                if (args.beta != 0.0) val += args.beta * (*dst);
                *dst = val;
            }
            if (jn + 1 < G::JN) {
                // every wave is done with column block jn: its slot takes region jn + 1 of
                // the next problem
                lds_barrier();
                load_b_share(next ? p + grid : p, next, jn + 1, slot_of(i + 1, jn + 1));
            }
        }
    }
}

// ---- streaming form (iters = 1): B sliced column block by column block ----
//
// Column block c of a problem needs only the digit planes of B rows 16c .. 16c + 15
// (region c), so B is streamed through LDS one region at a time, in a pipeline over the
// workgroup's consecutive regions t = JN k + c (problem k, block c): while the waves
// multiply region t, they slice region t + 1 and take the row maxima of region t + 2,
// whose raw rows were loaded D - 2 columns earlier; every wave works on its own share of
// each region (two 1 KB pieces: K step h = wave / 4 and four consecutive k of every
// lane's fragment slot group, so the slicing of a share writes one dword of each plane
// fragment).  The row exponent of a region is the maximum of the waves' partial row
// maxima (one LDS round through a barrier).  VALU slicing and MFMA products of the same
// column run in the same stretch of instructions, so they overlap.  A is sliced into
// registers at the first column of each problem; the next problem's A rows arrive in
// registers two or three 16-byte pieces per column over columns 0 .. JN - 2, so every
// column issues about the same number of loads and the memory system sees an even
// stream (a whole problem's A at once is 128 KB per workgroup, and vector memory
// operations complete in order, so the region waits of the next columns would wait for
// all of it).  LDS: D raw regions, two plane regions, partial maxima and exponents
// (~100 KB for 128^3 with D = 4).
template <class G, int D_>
struct I8Stream {
    static constexpr int D = D_;  // raw regions held: t + 1 .. t + D
    static constexpr int RAW_OFF = 0;
    static constexpr int PLN_OFF = D * G::REGION;
    static constexpr int PMAX_OFF = PLN_OFF + 2 * G::REGION;     // int [2][WAVES][16]
    static constexpr int EXPB_OFF = PMAX_OFF + 2 * G::WAVES * 16 * 4;  // int [2][16]
    static constexpr int EXPA_OFF = EXPB_OFF + 2 * 16 * 4;       // int [BM]
    static constexpr int LDS = EXPA_OFF + G::BM * 4;
    // a wave's share of a region is pieces 2 wave, 2 wave + 1 of the region's 8 KS
    static_assert(G::PPW == 2 && G::BM == G::BK, "two pieces per wave and region");
    static_assert(D >= 3, "a region's raw rows land at least one column before its maxima");
    static_assert(LDS <= 160 * 1024, "LDS budget");
    // The lane's 8 KS pieces of the next problem's A rows, spread over columns
    // 0 .. JN - 2 (piece pc: K step pc / 8, fragment slot group pc % 8): at least two
    // columns pass between a piece's load and its use at the next problem's column 0.
    static constexpr int PIECES = 8 * G::KS;
    static_assert(G::JN >= 3, "A pieces spread over at least two columns");
    static constexpr int a_count(int c) {
        return c >= G::JN - 1 ? 0 : PIECES / (G::JN - 1) + (c < PIECES % (G::JN - 1) ? 1 : 0);
    }
    static constexpr int a_first(int c) {
        int s = 0;
        for (int i = 0; i < c; ++i) s += a_count(i);
        return s;
    }
    static_assert(a_first(G::JN) == PIECES, "every piece is loaded once");
    // Column c issues, in order: its region load (2 operations), a_count(c) A loads and 4
    // stores.  Region t + 2 was loaded first thing at column t + 2 - D, so at column t it
    // has landed once at most the operations issued after it are outstanding.
    static constexpr int col(int c) { return (c % G::JN + G::JN) % G::JN; }
    static constexpr int region_wait(int c) {
        int n = a_count(col(c + 2 - D)) + 4;
        for (int j = c + 3 - D; j < c; ++j) n += 2 + a_count(col(j)) + 4;
        return n;
    }
    static_assert(region_wait(0) < 64 && region_wait(1) < 64 && region_wait(2) < 64, "vmcnt range");
};

template <class F, int... I>
__device__ inline void unroll_seq(F&& f, std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, I>{}), ...);
}

template <class G, int D, bool BETA>
__global__ void __launch_bounds__(G::NT)
split_i8_stream_kernel(SplitArgs args, int64_t count) {
    using St = I8Stream<G, D>;
    __shared__ __attribute__((aligned(16))) unsigned char smem[St::LDS];
    int* const pmax = reinterpret_cast<int*>(smem + St::PMAX_OFF);
    int* const expb = reinterpret_cast<int*>(smem + St::EXPB_OFF);
    int* const expa = reinterpret_cast<int*>(smem + St::EXPA_OFF);

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int r = lane & 15, grp = lane >> 4;
    const int row = wave * 16 + r;
    const int64_t first = blockIdx.x, grid = gridDim.x;
    if (first >= count) return;  // uniform
    const int64_t nprob = (count - first + grid - 1) / grid;  // this workgroup's problems
    // Per-lane byte offsets live in one VGPR each; everything wave-uniform (a region's rows,
    // a piece's K position, an output row) goes to the instructions' scalar offset or
    // immediate fields, so the unrolled columns add no per-column address registers.
    const uint32_t a_off = (uint32_t)((row * args.lda + 2 * grp) * 8);
    const uint32_t c_off = (uint32_t)(((wave * 16 + 4 * grp) * args.ldc + r) * 8);
    const int c_row = (int)(args.ldc * 8);
    // the wave's share: K step sh, pieces si and si + 1 (k = 64 sh + 8 (si + q) + 2 grp + {0, 1})
    const int sh = (2 * wave) >> 3, si = (2 * wave) & 7;
    const uint32_t b_lane = (uint32_t)((r * args.ldb + 2 * grp) * 8);
    const int b_share = (64 * sh + 8 * si) * 8;
    const int b_rows16 = (int)(16 * args.ldb * 8);

    // Buffer descriptors of problem q's operands; with `live` false a descriptor has no
    // range and its loads return zeros without touching memory.
    auto rsrc = [&](const double* base, int64_t q, int64_t stride, bool live) {
        return __builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(base + q * stride), (short)0,
                                                 live ? 0x7FFFFFF0 : 0, 0x00020000);
    };
    u4_t xa[G::KS][8];
    auto load_a_piece = [&](__amdgpu_buffer_rsrc_t rs, auto pc_tag) {
        constexpr int pc = decltype(pc_tag)::value, h = pc >> 3, i = pc & 7;
        xa[h][i] = __builtin_bit_cast(
            u4_t, __builtin_amdgcn_raw_buffer_load_b128(rs, (int)a_off, 8 * (64 * h + 8 * i), 0));
    };
    // Region t sits in raw slot t mod D.  With kb = (JN k) mod D the slot of region JN k
    // (wave-uniform; always 0 when JN is a multiple of D), region JN k + cc of problem k
    // (cc >= JN: region cc - JN of problem k + 1) takes slot (kb + cc) mod D.
    auto raw_slot = [&](int kb, int cc) {
        return (G::JN % St::D == 0 ? cc : kb + cc) % St::D;
    };
    // this wave's share of region JN k + cc into its raw slot
    auto dma = [&](int64_t k, int64_t p, int kb, auto cc_tag) {
        constexpr int cc = decltype(cc_tag)::value, dk = cc / G::JN, jn = cc % G::JN;
        const bool live = k + dk < nprob;
        const __amdgpu_buffer_rsrc_t rs = rsrc(args.B, live ? p + dk * grid : first, args.sb, live);
        unsigned char* dst = smem + St::RAW_OFF + raw_slot(kb, cc) * G::REGION + 2 * wave * G::FRAG;
        __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (__attribute__((address_space(3))) void*)dst, 16,
                                                 (int)b_lane, jn * b_rows16 + b_share, 0, 0);
        __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (__attribute__((address_space(3))) void*)(dst + G::FRAG),
                                                 16, (int)b_lane, jn * b_rows16 + b_share + 64, 0, 0);
    };
    auto raw_share = [&](int kb, auto cc_tag, u4_t (&x)[2]) {
        constexpr int cc = decltype(cc_tag)::value;
        const unsigned char* src =
            smem + St::RAW_OFF + raw_slot(kb, cc) * G::REGION + 2 * wave * G::FRAG + 16 * lane;
        x[0] = load_raw<16>(src);
        x[1] = load_raw<16>(src + G::FRAG);
    };
    // partial row maxima of the share (as frexp exponents: the exponent of a maximum is
    // the maximum of the exponents), into buffer `buf`
    auto pmax_write = [&](int buf, const u4_t (&x)[2]) {
        const d2_t x0 = bitcast<d2_t>(x[0]), x1 = bitcast<d2_t>(x[1]);
        double m = __builtin_fmax(__builtin_fmax(__builtin_fabs(x0[0]), __builtin_fabs(x0[1])),
                                  __builtin_fmax(__builtin_fabs(x1[0]), __builtin_fabs(x1[1])));
        m = __builtin_fmax(m, __shfl_xor(m, 16, 64));
        m = __builtin_fmax(m, __shfl_xor(m, 32, 64));
        pmax[(buf * G::WAVES + wave) * 16 + r] = __builtin_amdgcn_frexp_exp(m);
    };
    // digits of the share into plane slot reg mod 2: dword si / 2 of every lane's
    // fragment of plane p, K step sh
    auto slice_share = [&](int64_t reg, const u4_t (&x)[2]) {
        const int* pm = pmax + (int)(reg & 1) * G::WAVES * 16 + r;
        int E = pm[0];
#pragma unroll
        for (int w = 1; w < G::WAVES; ++w) E = max(E, pm[w * 16]);
        if (wave == 0 && grp == 0) expb[(int)(reg & 1) * 16 + r] = E;
        d2_t x0 = bitcast<d2_t>(x[0]), x1 = bitcast<d2_t>(x[1]);
        const int sc = E < kTinyRowExp ? 1000 : 0;  // tiny rows: scaled by 2^1000
        x0[0] = __builtin_amdgcn_ldexp(x0[0], sc);
        x0[1] = __builtin_amdgcn_ldexp(x0[1], sc);
        x1[0] = __builtin_amdgcn_ldexp(x1[0], sc);
        x1[1] = __builtin_amdgcn_ldexp(x1[1], sc);
        const uint32_t c_e = (uint32_t)(1023 + E + sc);
        uint32_t out[8];
        transpose4(digit_word(x0[0], c_e), digit_word(x0[1], c_e), digit_word(x1[0], c_e),
                   digit_word(x1[1], c_e), out);
        unsigned char* dst = smem + St::PLN_OFF + (int)(reg & 1) * G::REGION + sh * G::FRAG +
                             16 * lane + 4 * (si >> 1);
#pragma unroll
        for (int b = 0; b < 8; ++b)
            *reinterpret_cast<uint32_t*>(dst + (7 - b) * G::KS * G::FRAG) = out[b];
    };

    // ---- prologue: regions 0 .. D-1 and the first A rows in flight; maxima of regions
    //      0 and 1, digits of region 0 ----
    static_assert(G::JN % 2 == 0, "plane and maxima buffers fixed by the column index");
    unroll_seq([&](auto d) { dma((int64_t)0, first, 0, d); }, std::make_integer_sequence<int, St::D>{});
    {
        const __amdgpu_buffer_rsrc_t rs = rsrc(args.A, first, args.sa, true);
        unroll_seq([&](auto pc) { load_a_piece(rs, pc); }, std::make_integer_sequence<int, St::PIECES>{});
    }
    __builtin_amdgcn_s_waitcnt(vmcnt_only(0));
    {
        u4_t x[2];
        raw_share(0, std::integral_constant<int, 0>{}, x);
        pmax_write(0, x);
        raw_share(0, std::integral_constant<int, 1>{}, x);
        pmax_write(1, x);
        lds_barrier();
        raw_share(0, std::integral_constant<int, 0>{}, x);
        slice_share(0, x);
        lds_barrier();
    }

    u4_t ad[G::S][G::KS];
    int erow[4];
    int kb = 0;  // raw slot of the problem's region 0
    for (int64_t k = 0; k < nprob; ++k, kb = (kb + G::JN) % St::D) {
        const int64_t p = first + k * grid;
        const bool more = k + 1 < nprob;
        const __amdgpu_buffer_rsrc_t crs = rsrc(args.C, p, args.sc, true);
        // column block c of problem k (region t = JN k + c), unrolled over c: the plane and
        // maxima buffers (JN is even), the A pieces of a column and its wait count are fixed
        auto column_block = [&](auto c_tag) {
        constexpr int c = decltype(c_tag)::value;
        if constexpr (c == 0) {
            // ---- a new problem: A exponents and digits into registers ----
            const int ea = row_exponent<G::KS>(xa);
            expa[row] = ea;
            const bool tiny = __builtin_amdgcn_ballot_w64(ea < kTinyRowExp) != 0;
#pragma unroll
            for (int h = 0; h < G::KS; ++h) {
                u4_t f[8];
                slice_step(xa[h], ea, tiny, f);
#pragma unroll
                for (int q = 0; q < G::S; ++q) ad[q][h] = f[q];
                __builtin_amdgcn_sched_barrier(0);
            }
#pragma unroll
            for (int v = 0; v < 4; ++v) erow[v] = expa[wave * 16 + 4 * grp + v];  // own rows
        }
        // ---- B pipeline: maxima of region t + 2, digits of region t + 1 ----
        // this wave's loads of region t + 2 (issued two columns ago) have landed
        if constexpr (BETA)
            __builtin_amdgcn_s_waitcnt(vmcnt_only(0));  // C loads change the counts
        else
            __builtin_amdgcn_s_waitcnt(vmcnt_only(St::region_wait(c)));
        dma(k, p, kb, std::integral_constant<int, c + St::D>{});  // into the slot region t left
        // this column's share of the next problem's A pieces (the registers of problem k's
        // A were sliced at column 0)
        if constexpr (St::a_count(c) > 0) {
            const __amdgpu_buffer_rsrc_t rs = rsrc(args.A, more ? p + grid : p, args.sa, more);
            unroll_seq([&](auto j) { load_a_piece(rs, std::integral_constant<int, St::a_first(c) + j>{}); },
                       std::make_integer_sequence<int, St::a_count(c)>{});
        }
        u4_t x2[2], x1[2];
        raw_share(kb, std::integral_constant<int, c + 2>{}, x2);
        raw_share(kb, std::integral_constant<int, c + 1>{}, x1);
        // row exponents of region t + 1: the largest of the waves' partial maxima
        int E;
        {
            const int* pm = pmax + ((c + 1) & 1) * G::WAVES * 16 + r;
            E = pm[0];
#pragma unroll
            for (int w = 1; w < G::WAVES; ++w) E = max(E, pm[w * 16]);
        }
        // every wave and lane group writes the same value; written at column t, read at
        // column t + 1, so two buffers alternate across the barrier between them
        expb[((c + 1) & 1) * 16 + r] = E;
        const int ecol = expb[(c & 1) * 16 + r];
        const unsigned char* const bsrc = smem + St::PLN_OFF + (c & 1) * G::REGION + 16 * lane;
        auto bfrag = [&](int pb, int h) {
            return bitcast<i4_t>(load_raw<16>(bsrc + (pb * G::KS + h) * G::FRAG));
        };
        // The level products of region t (MFMA) and the VALU work of the column -- the
        // partial maxima of region t + 2 and the digits of region t + 1, independent of the
        // products -- in one stretch, interleaved in program order: the products run in
        // groups, one B fragment (K step h, plane pb) per group, each followed by one step
        // of the VALU work, so that the VALU issues while the MFMAs run.  Planes go in the
        // order 0, 7, 1, 6, 2, 5, 3, 4 (8, 1, 7, 2, 6, 3, 5, 4 products): any two
        // consecutive groups hold at least 9 products, and fragments are read two groups
        // ahead.  Rows of region t + 1 with E < -960 take the scaled slicing (a scale of 2^0
        // for the others: four ldexp per lane).
        static_assert(G::L == G::S && G::S == 8, "eight planes, all levels below 8");
        constexpr int kOrder[8] = {0, 7, 1, 6, 2, 5, 3, 4};
        constexpr int NF = 8 * G::KS;
        i4_t T[G::L];
#pragma unroll
        for (int l = 0; l < G::L; ++l) T[l] = i4_t{0, 0, 0, 0};
        i4_t fb[3];
        fb[0] = bfrag(kOrder[0], 0);
        fb[1] = bfrag(kOrder[1], 0);
        d2_t a0, a1;
        uint32_t c_e = 0;
        uint64_t dw[4];
        uint32_t out[8];
        auto valu_step = [&](auto s_tag) {
            constexpr int s = decltype(s_tag)::value;
            if constexpr (s == 0) {
                pmax_write(c & 1, x2);
            } else if constexpr (s == 1) {
                a0 = bitcast<d2_t>(x1[0]);
                a1 = bitcast<d2_t>(x1[1]);
                const int sc = E < kTinyRowExp ? 1000 : 0;
                a0[0] = __builtin_amdgcn_ldexp(a0[0], sc);
                a0[1] = __builtin_amdgcn_ldexp(a0[1], sc);
                a1[0] = __builtin_amdgcn_ldexp(a1[0], sc);
                a1[1] = __builtin_amdgcn_ldexp(a1[1], sc);
                c_e = (uint32_t)(1023 + E + sc);
            } else if constexpr (s >= 2 && s <= 5) {
                const double x = s == 2 ? a0[0] : s == 3 ? a0[1] : s == 4 ? a1[0] : a1[1];
                dw[s - 2] = digit_word(x, c_e);
            } else if constexpr (s == 6) {
                transpose4(dw[0], dw[1], dw[2], dw[3], out);
            } else if constexpr (s == 7) {
                unsigned char* dst = smem + St::PLN_OFF + ((c + 1) & 1) * G::REGION + sh * G::FRAG +
                                     16 * lane + 4 * (si >> 1);
#pragma unroll
                for (int b = 0; b < 8; ++b)
                    *reinterpret_cast<uint32_t*>(dst + (7 - b) * G::KS * G::FRAG) = out[b];
            }
        };
        unroll_seq(
            [&](auto f_tag) {
                constexpr int f = decltype(f_tag)::value, h = f / 8, pb = kOrder[f % 8];
                if constexpr (f + 2 < NF) fb[(f + 2) % 3] = bfrag(kOrder[(f + 2) % 8], (f + 2) / 8);
#pragma unroll
                for (int pa = 0; pa + pb < G::L; ++pa)
                    T[pa + pb] = __builtin_amdgcn_mfma_i32_16x16x64_i8(bitcast<i4_t>(ad[pa][h]), fb[f % 3],
                                                                       T[pa + pb], 0, 0, 0);
                // the VALU steps after every KS-th group: eight steps per column
                if constexpr (f % G::KS == 0) valu_step(std::integral_constant<int, f / G::KS>{});
                __builtin_amdgcn_sched_barrier(0);
            },
            std::make_integer_sequence<int, NF>{});
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            double acc = (double)T[G::L - 1][v];
#pragma unroll
            for (int l = G::L - 2; l >= 0; --l) acc = __builtin_fma(acc, 1.0 / 256.0, (double)T[l][v]);
            const int so = v * c_row + c * 128;
            double val = __builtin_amdgcn_ldexp(acc, erow[v] + ecol - 12) * args.alpha;
            if constexpr (BETA)
                val += args.beta * __builtin_bit_cast(
                                       double, __builtin_amdgcn_raw_buffer_load_b64(crs, (int)c_off, so, 0));
            __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2_t, val), crs, (int)c_off, so, 0);
        }
        lds_barrier();
        };
        unroll_seq(column_block, std::make_integer_sequence<int, G::JN>{});
    }
}

// Raw regions of B held in LDS by the streaming form: region loads run D - 2 columns
// ahead of their row maxima.
constexpr int kI8StreamDepth = 4;

// BETA: C is read (beta != 0); a separate instance, so that the columns of the other one
// carry no branch
template <class G, bool BETA>
hipError_t launch_split_i8_stream_form(const SplitArgs& a, int64_t count, hipStream_t s) {
    constexpr int D = kI8StreamDepth;
    static std::atomic<int> cache[kMaxCachedDevices];
    int slots = 0;
    const hipError_t e = resident_workgroups(split_i8_stream_kernel<G, D, BETA>, G::NT, cache, &slots);
    if (e != hipSuccess) return e;
    const int64_t grid = std::min<int64_t>(count, (int64_t)slots);
    hipLaunchKernelGGL((split_i8_stream_kernel<G, D, BETA>), dim3((unsigned)grid), dim3(G::NT), 0, s, a,
                       count);
    return hipGetLastError();
}

template <class G>
hipError_t launch_split_i8_stream(const SplitArgs& a, int64_t count, hipStream_t s) {
    return a.beta != 0.0 ? launch_split_i8_stream_form<G, true>(a, count, s)
                         : launch_split_i8_stream_form<G, false>(a, count, s);
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
                       : launch_split_i8_stream<G>(a, count, s);
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
