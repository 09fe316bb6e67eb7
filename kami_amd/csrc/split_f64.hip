// FP64 block GEMM through FP16 / BF16 slices on MFMA (split precision).
//
// C = A * B^T in FP64 computed as an error-free sum of low-precision slice products
// (the Ozaki scheme: Ozaki, Ogita, Oishi, Rump, "Error-free transformations of matrix
// multiplication by using fast routines of matrix multiplication and its applications",
// Numer. Algorithms 59 (2012); tensor-core form: Mukunoki, Ozaki, Ogita, Imamura,
// ICCS 2020).  No counterpart exists in the reference; it is the split-precision FP64
// path the block_gemm interface offers beside native FP64 MFMA.
//
// Per problem (one workgroup):
//   1. exponents: e_i = exponent of max_k |A(i,k)| for every row of A and f_j of every
//      row of B (B is stored N x K), so |A(i,k)| < 2^e_i;
//   2. slicing, per 32-wide K chunk: A(i,k) 2^(BETA - e_i) = sum_s a_s 2^(-BETA (s-1))
//      with integer digits a_s chosen by round-to-nearest (|a_1| <= 2^BETA,
//      |a_s| <= 2^(BETA-1) after that), exactly representable in FP16 (BETA = 9) or
//      BF16 (BETA = 8); the digits go to LDS, one plane per slice;
//   3. products: for every level g = s + t <= G the FP32 MFMA sum
//      T_g = sum_{s+t=g} a_s b_t^T over the 32-wide chunk is an integer of magnitude
//      <= 2^24 (checked at compile time for the levels that matter), hence exact in
//      any summation order; it is drained into an FP64 accumulator with weight
//      2^(-BETA g) after each chunk;
//   4. C(i,j) = alpha 2^(e_i + f_j) acc(i,j) + beta C(i,j).
// The error relative to 2^(e_i + f_j) is about K 2^(-BETA S - 1) from the truncated
// digits plus the dropped levels g > G; finite inputs are required.
#include <cmath>
#include <cstdint>

#include "split_common.hpp"

namespace kami {

// Largest |digit| sum bound of level g over KC terms, in units of 2^0, compared against
// 2^24 (the FP32 integer range).  The first digit has b1 bits (|d_1| <= 2^b1), the later
// ones BETA bits with round to nearest (|d_s| <= 2^(BETA-1)).
constexpr double level_bound(int beta, int S, int g, int kc, int b1) {
    double tot = 0;
    for (int s = 1; s <= S; ++s) {
        const int t = g - s;
        if (t < 1 || t > S) continue;
        const double bs = (s == 1) ? (double)(1 << b1) : (double)(1 << (beta - 1));
        const double bt = (t == 1) ? (double)(1 << b1) : (double)(1 << (beta - 1));
        tot += kc * bs * bt;
    }
    return tot;
}
constexpr double level_bound(int beta, int S, int g, int kc) { return level_bound(beta, S, g, kc, beta); }

// Every level whose FP32 rounding could matter is exact: a level that may exceed 2^24
// is allowed only when its worst rounding, 2^-24 * bound * 2^(-beta g + 2 (beta - b1)) in
// units of 2^(e_i + f_j), stays below 2^-50.
constexpr bool levels_exact_range(int beta, int S, int g0, int G, int kc, int b1) {
    for (int g = g0; g <= G; ++g) {
        const double b = level_bound(beta, S, g, kc, b1);
        if (b <= 16777216.0) continue;
        double rel = b;
        for (int q = 0; q < 24 + beta * g - 2 * (beta - b1); ++q) rel *= 0.5;
        double lim = 1.0;
        for (int q = 0; q < 50; ++q) lim *= 0.5;
        if (rel > lim) return false;
    }
    return true;
}
constexpr bool levels_exact(int beta, int S, int G, int kc) {
    return levels_exact_range(beta, S, 2, G, kc, beta);
}

// One level's part of levels_exact: its sum over kc terms is exact in FP32, or its worst
// rounding is below 2^-50 in units of 2^(e_i + f_j).
constexpr bool levels_exact_one(int beta, int S, int g, int kc, int b1) {
    return levels_exact_range(beta, S, g, g, kc, b1);
}

// 2^e as a compile-time constant (exact for the exponents used here).
constexpr double pow2(int e) {
    double r = 1.0;
    for (; e > 0; --e) r *= 2.0;
    for (; e < 0; ++e) r *= 0.5;
    return r;
}

template <class SAT, int BM_, int BN_, int BK_, int PM_, int PN_, int BETA_, int S_, int G_>
struct SplitGeom {
    static constexpr int BM = BM_, BN = BN_, BK = BK_, PM = PM_, PN = PN_;
    static constexpr int BETA = BETA_, S = S_, G = G_;
    static constexpr int WAVES = PM * PN, NT = WAVES * kWave;
    static constexpr int TM = BM / PM, TN = BN / PN, IM = TM / 16, JN = TN / 16;
    static constexpr int JH = JN < 2 ? JN : 2;  // column atoms per register group
    static constexpr int KC = 32, NCH = BK / KC;
    // 32 digits per row; the conflict-free stride when S digit planes of it fit in LDS,
    // else 16 bytes of padding
    static constexpr int ROWB_CF = lds_row_bytes(KC * 2, 16);
    static constexpr int ROWB =
        S * (BM + BN) * ROWB_CF + (BM + BN) * 4 <= 160 * 1024 ? ROWB_CF : KC * 2 + 16;
    static constexpr int SA_BYTES = BM * ROWB, SB_BYTES = BN * ROWB;
    static constexpr int PLANE = SA_BYTES + SB_BYTES;
    static constexpr int EXP_OFF = S * PLANE;
    static constexpr int LDS = EXP_OFF + (BM + BN) * 4;
    static constexpr int PA = BM * 16 / NT, PB = BN * 16 / NT;  // 16-byte pieces per thread
    static_assert(BM % PM == 0 && BN % PN == 0 && TM % 16 == 0 && TN % 16 == 0, "tiling");
    static_assert(BK % KC == 0, "K must be a multiple of 32");
    static_assert((BM * 16) % NT == 0 && (BN * 16) % NT == 0, "piece distribution");
    static_assert(LDS <= 160 * 1024, "LDS budget");
    static_assert(levels_exact(BETA, S, G, KC), "slice digits would not sum exactly in FP32");
    // the persistent form adds the two smallest levels in FP32 before converting them
    static_assert(G >= 3 && level_bound(BETA, S, G - 1, KC) * pow2(-24 - BETA * (G - 1)) <= pow2(-50),
                  "the two smallest levels are combined in FP32");
};

__device__ inline int exp_of(double x) {
    // exponent e with |x| in [2^(e-1), 2^e); zero maps to a sentinel below every
    // finite exponent so it never raises a row maximum
    return x == 0.0 ? -4096 : __builtin_amdgcn_frexp_exp(x);
}

// Maximum over the 16 lanes of a DPP row (lanes 16r .. 16r + 15), by lane permutations
// within the row: swap neighbours, swap pairs, mirror each half row, mirror the row.
// An all-zero row (sentinel exponent) maps to 0.
__device__ inline int row16_max_exp(int e) {
    e = max(e, __builtin_amdgcn_update_dpp(0, e, 0xB1, 0xF, 0xF, false));   // quad_perm 1,0,3,2
    e = max(e, __builtin_amdgcn_update_dpp(0, e, 0x4E, 0xF, 0xF, false));   // quad_perm 2,3,0,1
    e = max(e, __builtin_amdgcn_update_dpp(0, e, 0x141, 0xF, 0xF, false));  // row_half_mirror
    e = max(e, __builtin_amdgcn_update_dpp(0, e, 0x140, 0xF, 0xF, false));  // row_mirror
    return e == -4096 ? 0 : e;
}

template <class SAT>
__device__ inline uint32_t pack2(double a, double b);

template <>
__device__ inline uint32_t pack2<AtomF16x32>(double a, double b) {
    // Small integers: exact in float and in half.  The conversion goes through float
    // explicitly (f64 -> f32, then one packed f32 -> f16 pair conversion, exact here): a
    // direct double -> half conversion has no instruction and is emulated in integer code.
    const float fa = (float)a, fb = (float)b;
    return __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pkrtz(fa, fb));
}

template <>
__device__ inline uint32_t pack2<AtomBF16x32>(double a, double b) {
    // |digit| <= 256: exact in float and in bfloat16 (upper half of the float)
    const uint32_t ua = __builtin_bit_cast(uint32_t, (float)a);
    const uint32_t ub = __builtin_bit_cast(uint32_t, (float)b);
    return (ua >> 16) | (ub & 0xffff0000u);
}

template <class SAT, class G>
__global__ void __launch_bounds__(G::NT)
split_gemm_kernel(SplitArgs args) {
    using frag_t = typename SAT::frag_t;
    using accv_t = typename SAT::accv_t;
    __shared__ __attribute__((aligned(16))) unsigned char smem[G::LDS];
    int* const expo = reinterpret_cast<int*>(smem + G::EXP_OFF);  // [BM] then [BN]

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int pn = wave % G::PN, pm = wave / G::PN;
    const int64_t prob = blockIdx.x;
    const double* __restrict__ A = args.A + prob * args.sa;
    const double* __restrict__ B = args.B + prob * args.sb;
    double* __restrict__ C = args.C + prob * args.sc;
    const int64_t lda = args.lda, ldb = args.ldb, ldc = args.ldc;

    // ---- 1. row exponents of A and B.  Thread t owns the 16-byte pieces
    //      p = t + i * NT of every K chunk (row p / 16, k pair p % 16), so the 16 lanes
    //      of a row reduce within one DPP row; every chunk's loads are independent. ----
    int ea[G::PA], eb[G::PB];
#pragma unroll
    for (int i = 0; i < G::PA; ++i) ea[i] = -4096;
#pragma unroll
    for (int i = 0; i < G::PB; ++i) eb[i] = -4096;
#pragma unroll
    for (int c = 0; c < G::NCH; ++c) {
#pragma unroll
        for (int i = 0; i < G::PA; ++i) {
            const int p = tid + i * G::NT, row = p >> 4, kp = p & 15;
            const d2_t x = bitcast<d2_t>(load_raw<16>(A + row * lda + c * G::KC + 2 * kp));
            ea[i] = max(ea[i], max(exp_of(x[0]), exp_of(x[1])));
        }
#pragma unroll
        for (int i = 0; i < G::PB; ++i) {
            const int p = tid + i * G::NT, row = p >> 4, kp = p & 15;
            const d2_t x = bitcast<d2_t>(load_raw<16>(B + row * ldb + c * G::KC + 2 * kp));
            eb[i] = max(eb[i], max(exp_of(x[0]), exp_of(x[1])));
        }
    }
#pragma unroll
    for (int i = 0; i < G::PA; ++i) ea[i] = row16_max_exp(ea[i]);
#pragma unroll
    for (int i = 0; i < G::PB; ++i) eb[i] = row16_max_exp(eb[i]);
#pragma unroll
    for (int i = 0; i < G::PA; ++i) {
        const int p = tid + i * G::NT;
        if ((p & 15) == 0) expo[p >> 4] = ea[i];
    }
#pragma unroll
    for (int i = 0; i < G::PB; ++i) {
        const int p = tid + i * G::NT;
        if ((p & 15) == 0) expo[G::BM + (p >> 4)] = eb[i];
    }

    double acc[G::IM][G::JN][4];
#pragma unroll
    for (int i = 0; i < G::IM; ++i)
#pragma unroll
        for (int j = 0; j < G::JN; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[i][j][v] = 0.0;

    // ---- chunk pieces: 16 bytes = 2 doubles; 16 consecutive threads cover one row ----
    u4_t xa[G::PA], xb[G::PB];
    auto load_chunk = [&](int c) {
#pragma unroll
        for (int i = 0; i < G::PA; ++i) {
            const int p = tid + i * G::NT, row = p >> 4, kp = p & 15;
            xa[i] = load_raw<16>(A + row * lda + c * G::KC + 2 * kp);
        }
#pragma unroll
        for (int i = 0; i < G::PB; ++i) {
            const int p = tid + i * G::NT, row = p >> 4, kp = p & 15;
            xb[i] = load_raw<16>(B + row * ldb + c * G::KC + 2 * kp);
        }
    };
    auto slice_piece = [&](const u4_t& raw, int row, int kp, int e, unsigned char* plane0) {
        const d2_t x = bitcast<d2_t>(raw);
        double r0 = __builtin_amdgcn_ldexp(x[0], G::BETA - e);
        double r1 = __builtin_amdgcn_ldexp(x[1], G::BETA - e);
#pragma unroll
        for (int s = 0; s < G::S; ++s) {
            const double q0 = __builtin_rint(r0), q1 = __builtin_rint(r1);
            *reinterpret_cast<uint32_t*>(plane0 + s * G::PLANE + row * G::ROWB + kp * 4) =
                pack2<SAT>(q0, q1);
            if (s + 1 < G::S) {
                r0 = __builtin_amdgcn_ldexp(r0 - q0, G::BETA);
                r1 = __builtin_amdgcn_ldexp(r1 - q1, G::BETA);
            }
        }
    };

    load_chunk(0);
    __syncthreads();  // exponents visible

    for (int c = 0; c < G::NCH; ++c) {
        // ---- 2. slice the chunk into digit planes ----
#pragma unroll
        for (int i = 0; i < G::PA; ++i) {
            const int p = tid + i * G::NT, row = p >> 4, kp = p & 15;
            slice_piece(xa[i], row, kp, ea[i], smem);
        }
#pragma unroll
        for (int i = 0; i < G::PB; ++i) {
            const int p = tid + i * G::NT, row = p >> 4, kp = p & 15;
            slice_piece(xb[i], row, kp, eb[i], smem + G::SA_BYTES);
        }
        if (c + 1 < G::NCH) load_chunk(c + 1);  // in flight during the products
        __syncthreads();

        // ---- 3. level products on MFMA, drained to FP64 ----
        frag_t af[G::S][G::IM];
#pragma unroll
        for (int s = 0; s < G::S; ++s)
#pragma unroll
            for (int im = 0; im < G::IM; ++im)
                af[s][im] = bitcast<frag_t>(load_raw<16>(
                    smem + s * G::PLANE + (pm * G::TM + im * 16 + (lane & 15)) * G::ROWB +
                    16 * (lane >> 4)));
        // the wave tile is processed in column groups of JH atoms; a scheduling barrier
        // between groups keeps their fragments and level sums from being live together
#pragma unroll
        for (int h = 0; h < G::JN / G::JH; ++h) {
#pragma unroll
            for (int g = G::G; g >= 2; --g) {
                accv_t T[G::IM][G::JH];
#pragma unroll
                for (int im = 0; im < G::IM; ++im)
#pragma unroll
                    for (int jh = 0; jh < G::JH; ++jh) T[im][jh] = accv_t{0, 0, 0, 0};
#pragma unroll
                for (int s = 1; s <= G::S; ++s) {
                    const int t = g - s;
                    if (t < 1 || t > G::S) continue;
#pragma unroll
                    for (int jh = 0; jh < G::JH; ++jh) {
                        const int jn = h * G::JH + jh;
                        const frag_t bf = bitcast<frag_t>(load_raw<16>(
                            smem + G::SA_BYTES + (t - 1) * G::PLANE +
                            (pn * G::TN + jn * 16 + (lane & 15)) * G::ROWB + 16 * (lane >> 4)));
#pragma unroll
                        for (int im = 0; im < G::IM; ++im) SAT::mma(T[im][jh], af[s - 1][im], bf);
                    }
                }
                const double w = pow2(-G::BETA * g);
#pragma unroll
                for (int im = 0; im < G::IM; ++im)
#pragma unroll
                    for (int jh = 0; jh < G::JH; ++jh)
#pragma unroll
                        for (int v = 0; v < 4; ++v)
                            acc[im][h * G::JH + jh][v] =
                                __builtin_fma((double)T[im][jh][v], w, acc[im][h * G::JH + jh][v]);
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        __syncthreads();  // digit planes are rewritten by the next chunk
    }

    // ---- 4. rescale and store ----
#pragma unroll
    for (int im = 0; im < G::IM; ++im)
#pragma unroll
        for (int jn = 0; jn < G::JN; ++jn) {
            const int col = pn * G::TN + jn * 16 + (lane & 15);
            const int ecol = expo[G::BM + col];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = pm * G::TM + im * 16 + 4 * (lane >> 4) + v;
                double val = __builtin_amdgcn_ldexp(acc[im][jn][v], expo[row] + ecol);
                double* dst = C + row * ldc + col;
                val *= args.alpha;
                if (args.beta != 0.0) val += args.beta * (*dst);
                *dst = val;
            }
        }
}

// Persistent form.  A workgroup walks work units blockIdx.x, + gridDim.x, ... -- one unit
// is a BM x BN tile of C of one problem (the whole problem for a block GEMM; a tile of a
// larger GEMM for kami_split_gemm, with K any multiple of 32) -- and reads every operand
// once: the exponents are taken per row and per 32-wide K chunk (so no
// separate exponent pass over K is needed), the chunk is sliced from registers, and the
// next chunk -- chunk 0 of the next problem after the last one -- is loaded while the
// MFMAs of the current chunk run.  A chunk's level sums are combined in FP64 with the
// weights 2^(-BETA g) and then scaled by 2^(e_i + f_j) of that chunk into the
// accumulator, which holds C in absolute units.  Per-chunk exponents bound each chunk's
// digits by the chunk's own row maximum, so the truncation error is no larger than
// with exponents over the full K.
template <class SAT, class G>
__global__ void __launch_bounds__(G::NT)
split_pipe_kernel(SplitArgs args, int64_t count) {
    using frag_t = typename SAT::frag_t;
    using accv_t = typename SAT::accv_t;
    __shared__ __attribute__((aligned(16))) unsigned char smem[G::LDS];
    int* const expo = reinterpret_cast<int*>(smem + G::EXP_OFF);  // [BM] then [BN], per chunk

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int pn = wave % G::PN, pm = wave / G::PN;
    const int64_t grid = gridDim.x;
    const int64_t lda = args.lda, ldb = args.ldb, ldc = args.ldc;
    const int nch = args.nch;
    const int64_t tiles = (int64_t)args.tiles_m * args.tiles_n;
    // work unit u = one BM x BN tile of one problem; tiles of a problem are consecutive
    // with the column tile fastest, so neighbouring workgroups share A row panels
    auto unit_base = [&](int64_t u, int64_t& a_off, int64_t& b_off, int64_t& c_off) {
        const int64_t prob = u / tiles;
        const int t = (int)(u - prob * tiles);
        const int tm = t / args.tiles_n, tn = t - tm * args.tiles_n;
        a_off = prob * args.sa + (int64_t)tm * G::BM * lda;
        b_off = prob * args.sb + (int64_t)tn * G::BN * ldb;
        c_off = prob * args.sc + (int64_t)tm * G::BM * ldc + (int64_t)tn * G::BN;
    };

    // Slicing layout: thread t owns row t / 4 of A and of B and the 8 consecutive k values
    // 8 (t mod 4) .. +7 of each chunk (four 16-byte loads from one pointer; one 16-byte
    // LDS write per digit plane).  The pointers advance by one chunk per step and are set
    // up once per work unit.
    static_assert(G::NT == 4 * G::BM && G::NT == 4 * G::BN, "four threads per operand row");
    const int srow = tid >> 2, sk = 8 * (tid & 3);
    const double* pa;
    const double* pb;
    auto unit_ptrs = [&](int64_t u) {
        int64_t a_off, b_off, c_off;
        unit_base(u, a_off, b_off, c_off);
        pa = args.A + a_off + srow * lda + sk;
        pb = args.B + b_off + srow * ldb + sk;
    };
    u4_t xa[4], xb[4];
    auto load_chunk = [&]() {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            xa[i] = load_raw<16>(pa + 2 * i);
            xb[i] = load_raw<16>(pb + 2 * i);
        }
    };
    // Exponent e with |x| < 2^e over the row's 32 values of the chunk: the largest upper
    // word of |x| (monotonic in |x|) over the thread's 8 values and the row's 4 threads (one
    // DPP quad), completed with all-ones lower bits, so frexp also covers subnormal rows.
    auto chunk_exp = [&](const u4_t (&x)[4]) {
        uint32_t h = 0;
#pragma unroll
        for (int i = 0; i < 4; ++i) h = max(h, max(x[i][1] & 0x7fffffffu, x[i][3] & 0x7fffffffu));
        h = max(h, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)h, 0xB1, 0xF, 0xF, false));  // quad_perm 1,0,3,2
        h = max(h, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)h, 0x4E, 0xF, 0xF, false));  // quad_perm 2,3,0,1
        return __builtin_amdgcn_frexp_exp(__builtin_bit_cast(double, ((uint64_t)h << 32) | 0xffffffffu));
    };
    auto slice = [&](const u4_t (&x)[4], int e, unsigned char* plane0) {
        unsigned char* dst = plane0 + srow * G::ROWB + 2 * sk;
        double r[8];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const d2_t d = bitcast<d2_t>(x[i]);
            r[2 * i] = __builtin_amdgcn_ldexp(d[0], G::BETA - e);
            r[2 * i + 1] = __builtin_amdgcn_ldexp(d[1], G::BETA - e);
        }
#pragma unroll
        for (int s = 0; s < G::S; ++s) {
            u4_t packed;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const double q0 = __builtin_rint(r[2 * j]), q1 = __builtin_rint(r[2 * j + 1]);
                packed[j] = pack2<SAT>(q0, q1);
                if (s + 1 < G::S) {
                    r[2 * j] = (r[2 * j] - q0) * (double)(1 << G::BETA);
                    r[2 * j + 1] = (r[2 * j + 1] - q1) * (double)(1 << G::BETA);
                }
            }
            store_raw<16>(dst + s * G::PLANE, packed);
        }
    };

    int64_t p = blockIdx.x;  // work unit
    if (p >= count) return;  // uniform
    unit_ptrs(p);
    load_chunk();

    for (; p < count; p += grid) {
        double acc[G::IM][G::JN][4];
#pragma unroll
        for (int i = 0; i < G::IM; ++i)
#pragma unroll
            for (int j = 0; j < G::JN; ++j)
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[i][j][v] = 0.0;

        for (int c = 0; c < nch; ++c) {
            // ---- exponents and digit planes of this chunk ----
            const int ea = chunk_exp(xa), eb = chunk_exp(xb);
            if ((tid & 3) == 0) {
                expo[srow] = ea;
                expo[G::BM + srow] = eb;
            }
            slice(xa, ea, smem);
            slice(xb, eb, smem + G::SA_BYTES);
            // the next chunk is in flight during the products
            if (c + 1 < nch) {
                pa += G::KC;
                pb += G::KC;
                load_chunk();
            } else if (p + grid < count) {
                unit_ptrs(p + grid);
                load_chunk();
            }
            __syncthreads();

            // ---- level products on MFMA, combined in FP64, scaled into acc ----
            // One 16 x 16 atom at a time: every level's sum of the atom is live, the A
            // digit fragments of the atom row stay in registers across its columns, and
            // each B digit plane is read once per atom (plane t feeds levels t + 1 ...).
#pragma unroll
            for (int im = 0; im < G::IM; ++im) {
                frag_t af[G::S];
#pragma unroll
                for (int s = 0; s < G::S; ++s)
                    af[s] = bitcast<frag_t>(load_raw<16>(
                        smem + s * G::PLANE + (pm * G::TM + im * 16 + (lane & 15)) * G::ROWB +
                        16 * (lane >> 4)));
                int erow[4];
#pragma unroll
                for (int v = 0; v < 4; ++v) erow[v] = expo[pm * G::TM + im * 16 + 4 * (lane >> 4) + v];
#pragma unroll
                for (int jn = 0; jn < G::JN; ++jn) {
                    accv_t T[G::G - 1];  // level g in T[g - 2]
#pragma unroll
                    for (int l = 0; l < G::G - 1; ++l) T[l] = accv_t{0, 0, 0, 0};
#pragma unroll
                    for (int t = 1; t <= G::S; ++t) {
                        if (t + 1 > G::G) continue;
                        const frag_t bf = bitcast<frag_t>(load_raw<16>(
                            smem + G::SA_BYTES + (t - 1) * G::PLANE +
                            (pn * G::TN + jn * 16 + (lane & 15)) * G::ROWB + 16 * (lane >> 4)));
#pragma unroll
                        for (int s = 1; s <= G::S; ++s)
                            if (s + t <= G::G) SAT::mma(T[s + t - 2], af[s - 1], bf);
                    }
                    const int ecol = expo[G::BM + pn * G::TN + jn * 16 + (lane & 15)];
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        // Horner from the smallest weight, in units of level 2; the two
                        // smallest levels are added in FP32 first (rounding below 2^-50)
                        double cs = (double)__builtin_fmaf(T[G::G - 2][v], (float)pow2(-G::BETA),
                                                           T[G::G - 3][v]);
#pragma unroll
                        for (int g = G::G - 2; g >= 2; --g)
                            cs = __builtin_fma(cs, pow2(-G::BETA), (double)T[g - 2][v]);
                        acc[im][jn][v] += __builtin_amdgcn_ldexp(cs, erow[v] + ecol - 2 * G::BETA);
                    }
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
            __syncthreads();  // planes and exponents are rewritten by the next chunk
        }

        // ---- store ----
        int64_t a_off, b_off, c_off;
        unit_base(p, a_off, b_off, c_off);
        double* __restrict__ C = args.C + c_off;
#pragma unroll
        for (int im = 0; im < G::IM; ++im)
#pragma unroll
            for (int jn = 0; jn < G::JN; ++jn) {
                const int col = pn * G::TN + jn * 16 + (lane & 15);
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int row = pm * G::TM + im * 16 + 4 * (lane >> 4) + v;
                    double* dst = C + row * ldc + col;
                    double val = acc[im][jn][v] * args.alpha;
                    if (args.beta != 0.0) val += args.beta * (*dst);
                    *dst = val;
                }
            }
    }
}

// Resident-repeat geometry: the whole problem (BM x BK of A, BN x BK of B) is sliced
// once into LDS, S digit planes of full-K rows.  The first digit has B1 <= BETA bits
// (x 2^(B1 - e) = sum_s d_s 2^(-BETA (s-1))): one bit less than BETA keeps more level
// sums exact over the whole K, at one bit of the representation.
template <class SAT, int BM_, int BN_, int BK_, int PM_, int PN_, int BETA_, int S_, int G_,
          int B1_>
struct SplitResGeom {
    static constexpr int BM = BM_, BN = BN_, BK = BK_, PM = PM_, PN = PN_;
    static constexpr int BETA = BETA_, S = S_, G = G_, B1 = B1_;
    static constexpr int WAVES = PM * PN, NT = WAVES * kWave;
    static constexpr int TM = BM / PM, TN = BN / PN, IM = TM / 16, JN = TN / 16;
    static constexpr int KC = 32, NCH = BK / KC;
    static constexpr int ROWB = lds_row_bytes(BK * 2, 16);  // BK digits per row
    static constexpr int PLANE = (BM + BN) * ROWB;          // rows of A, then rows of B
    static constexpr int EXP_OFF = S * PLANE;
    static constexpr int LDS = EXP_OFF + (BM + BN) * 4;
    static constexpr int CPR = BK / 2;  // 16-byte pieces (two doubles) per row
    static constexpr int PA = BM * CPR / NT, PB = BN * CPR / NT;
    static_assert(BM % PM == 0 && BN % PN == 0 && TM % 16 == 0 && TN % 16 == 0, "tiling");
    static_assert(BK % KC == 0 && CPR <= 64 && 64 % CPR == 0, "a row's pieces sit in one wave");
    static_assert((BM * CPR) % NT == 0 && (BN * CPR) % NT == 0, "piece distribution");
    static_assert(LDS <= 160 * 1024, "LDS budget");
    static_assert(B1 <= BETA && levels_exact_range(BETA, S, 2, G, KC, B1),
                  "slice digits would not sum exactly in FP32");
    // weight of level g in units of 2^(e_i + f_j)
    static constexpr double weight(int g) { return pow2(-BETA * g + 2 * (BETA - B1)); }
    // Level g may run its FP32 sum over the whole K (drained once per repetition) when
    // that sum is exact or its worst rounding is below 2^-50 of 2^(e_i + f_j); the other
    // levels are drained after every 32-wide chunk.
    static constexpr bool whole_k(int g) { return levels_exact_one(BETA, S, g, BK, B1); }
    // The two smallest levels are combined in FP32 (T_(G-1) + T_G 2^-BETA) before their
    // one conversion; the rounding of that sum must stay below 2^-50 as well.
    static_assert(G >= 3 && whole_k(G) && whole_k(G - 1) &&
                      level_bound(BETA, S, G - 1, BK, B1) * pow2(-24) * weight(G - 1) <= pow2(-50),
                  "the two smallest levels are combined in FP32");
};

// Maximum exponent over the CPR consecutive lanes that hold one row (CPR = 16 or 32 here;
// any power of two up to 64), the all-zero sentinel mapped to 0 only after the full
// reduction.
template <int CPR>
__device__ inline int row_max_exp(int e) {
#pragma unroll
    for (int off = 1; off < CPR; off <<= 1) e = max(e, __shfl_xor(e, off, 64));
    return e == -4096 ? 0 : e;
}

// Resident-repeat form (`iters` > 1, the reference's NUM_ITER loop): one workgroup per
// problem.  The operands are read once, the exponents taken per row over the full K, and
// every digit plane of A and B written to LDS once; each wave then loads the digits of its
// rows of A and of B into registers, and the `iters` repetitions run only the level
// products and their FP64 recombination.  Level sums T_g that stay exact over the whole
// K (whole_k) are accumulated by the MFMAs across the chunks and added into the FP64
// accumulator once per repetition, the others after every 32-wide chunk; the two smallest
// levels are combined in FP32 first.  Each addition is acc += T_g weight(g).
template <class SAT, class G>
__global__ void __launch_bounds__(G::NT)
split_resident_kernel(SplitArgs args) {
    using frag_t = typename SAT::frag_t;
    using accv_t = typename SAT::accv_t;
    __shared__ __attribute__((aligned(16))) unsigned char smem[G::LDS];
    int* const expo = reinterpret_cast<int*>(smem + G::EXP_OFF);  // [BM] then [BN]

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int pn = wave % G::PN, pm = wave / G::PN;
    const int64_t prob = blockIdx.x;
    const double* __restrict__ A = args.A + prob * args.sa;
    const double* __restrict__ B = args.B + prob * args.sb;
    double* __restrict__ C = args.C + prob * args.sc;
    const int64_t lda = args.lda, ldb = args.ldb, ldc = args.ldc;

    // ---- operands: piece q = tid + i NT is row q / CPR, k pair q % CPR ----
    u4_t xa[G::PA], xb[G::PB];
#pragma unroll
    for (int i = 0; i < G::PA; ++i) {
        const int q = tid + i * G::NT;
        xa[i] = load_raw<16>(A + (q / G::CPR) * lda + 2 * (q % G::CPR));
    }
#pragma unroll
    for (int i = 0; i < G::PB; ++i) {
        const int q = tid + i * G::NT;
        xb[i] = load_raw<16>(B + (q / G::CPR) * ldb + 2 * (q % G::CPR));
    }
    auto slice_piece = [&](const u4_t& raw, int row, int kp, unsigned char* plane0) {
        const d2_t x = bitcast<d2_t>(raw);
        const int e = row_max_exp<G::CPR>(max(exp_of(x[0]), exp_of(x[1])));
        if (kp == 0) reinterpret_cast<int*>(smem + G::EXP_OFF)[row] = e;
        double r0 = __builtin_amdgcn_ldexp(x[0], G::B1 - e);
        double r1 = __builtin_amdgcn_ldexp(x[1], G::B1 - e);
#pragma unroll
        for (int s = 0; s < G::S; ++s) {
            const double q0 = __builtin_rint(r0), q1 = __builtin_rint(r1);
            *reinterpret_cast<uint32_t*>(plane0 + s * G::PLANE + row * G::ROWB + kp * 4) =
                pack2<SAT>(q0, q1);
            if (s + 1 < G::S) {
                r0 = (r0 - q0) * (double)(1 << G::BETA);
                r1 = (r1 - q1) * (double)(1 << G::BETA);
            }
        }
    };
#pragma unroll
    for (int i = 0; i < G::PA; ++i) {
        const int q = tid + i * G::NT;
        slice_piece(xa[i], q / G::CPR, q % G::CPR, smem);
    }
#pragma unroll
    for (int i = 0; i < G::PB; ++i) {
        const int q = tid + i * G::NT;
        slice_piece(xb[i], G::BM + q / G::CPR, q % G::CPR, smem);
    }
    __syncthreads();

    // digits of the wave's rows of A and B, every chunk and slice, held in registers
    frag_t af[G::NCH][G::S][G::IM], bfr[G::NCH][G::S][G::JN];
#pragma unroll
    for (int c = 0; c < G::NCH; ++c)
#pragma unroll
        for (int s = 0; s < G::S; ++s) {
#pragma unroll
            for (int im = 0; im < G::IM; ++im)
                af[c][s][im] = bitcast<frag_t>(load_raw<16>(
                    smem + s * G::PLANE + (pm * G::TM + im * 16 + (lane & 15)) * G::ROWB +
                    c * G::KC * 2 + 16 * (lane >> 4)));
#pragma unroll
            for (int jn = 0; jn < G::JN; ++jn)
                bfr[c][s][jn] = bitcast<frag_t>(load_raw<16>(
                    smem + s * G::PLANE + (G::BM + pn * G::TN + jn * 16 + (lane & 15)) * G::ROWB +
                    c * G::KC * 2 + 16 * (lane >> 4)));
        }

    double acc[G::IM][G::JN][4];
#pragma unroll
    for (int i = 0; i < G::IM; ++i)
#pragma unroll
        for (int j = 0; j < G::JN; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[i][j][v] = 0.0;

    for (int it = 0; it < args.iters; ++it) {
#pragma unroll
        for (int jn = 0; jn < G::JN; ++jn) {
            accv_t T[G::IM][G::G - 1];  // level g of atom row im in T[im][g - 2]
#pragma unroll
            for (int im = 0; im < G::IM; ++im)
#pragma unroll
                for (int l = 0; l < G::G - 1; ++l) T[im][l] = accv_t{0, 0, 0, 0};
#pragma unroll
            for (int c = 0; c < G::NCH; ++c) {
#pragma unroll
                for (int t = 1; t <= G::S; ++t) {
                    if (t + 1 > G::G) continue;
#pragma unroll
                    for (int s = 1; s <= G::S; ++s)
                        if (s + t <= G::G)
#pragma unroll
                            for (int im = 0; im < G::IM; ++im)
                                SAT::mma(T[im][s + t - 2], af[c][s - 1][im], bfr[c][t - 1][jn]);
                }
                // levels that are exact only within a chunk
#pragma unroll
                for (int g = G::G; g >= 2; --g) {
                    if (G::whole_k(g)) continue;
#pragma unroll
                    for (int im = 0; im < G::IM; ++im)
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            acc[im][jn][v] = __builtin_fma((double)T[im][g - 2][v],
                                                           G::weight(g), acc[im][jn][v]);
                            T[im][g - 2][v] = 0.0f;
                        }
                }
            }
            // whole-K levels, the two smallest combined in FP32 first
#pragma unroll
            for (int im = 0; im < G::IM; ++im)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const float low = __builtin_fmaf(T[im][G::G - 2][v], (float)pow2(-G::BETA),
                                                     T[im][G::G - 3][v]);
                    double a = __builtin_fma((double)low, G::weight(G::G - 1), acc[im][jn][v]);
#pragma unroll
                    for (int g = G::G - 2; g >= 2; --g)
                        if (G::whole_k(g)) a = __builtin_fma((double)T[im][g - 2][v], G::weight(g), a);
                    acc[im][jn][v] = a;
                }
        }
    }

    // ---- rescale and store ----
#pragma unroll
    for (int im = 0; im < G::IM; ++im)
#pragma unroll
        for (int jn = 0; jn < G::JN; ++jn) {
            const int col = pn * G::TN + jn * 16 + (lane & 15);
            const int ecol = expo[G::BM + col];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = pm * G::TM + im * 16 + 4 * (lane >> 4) + v;
                double val = __builtin_amdgcn_ldexp(acc[im][jn][v], expo[row] + ecol);
                double* dst = C + row * ldc + col;
                val *= args.alpha;
                if (args.beta != 0.0) val += args.beta * (*dst);
                *dst = val;
            }
        }
}

template <class SAT, class G>
hipError_t launch_split(const SplitArgs& a, int64_t count, hipStream_t s) {
    hipLaunchKernelGGL((split_gemm_kernel<SAT, G>), dim3((unsigned)count), dim3(G::NT), 0, s, a);
    return hipGetLastError();
}

// Persistent grid: every CU filled to the kernel's occupancy, never more workgroups
// than problems.
template <class SAT, class G>
hipError_t launch_split_pipe(const SplitArgs& a, int64_t count, hipStream_t s) {
    static std::atomic<int> cache[kMaxCachedDevices];
    int slots = 0;
    const hipError_t e = resident_workgroups(split_pipe_kernel<SAT, G>, G::NT, cache, &slots);
    if (e != hipSuccess) return e;
    const int64_t grid = std::min<int64_t>(count, (int64_t)slots);
    hipLaunchKernelGGL((split_pipe_kernel<SAT, G>), dim3((unsigned)grid), dim3(G::NT), 0, s, a,
                       count);
    return hipGetLastError();
}

template <class SAT, class G>
hipError_t launch_split_resident(const SplitArgs& a, int64_t count, hipStream_t s) {
    hipLaunchKernelGGL((split_resident_kernel<SAT, G>), dim3((unsigned)count), dim3(G::NT), 0, s,
                       a);
    return hipGetLastError();
}

template <class SAT, int PREC, int BM, int BN, int BK, int PM, int PN, int BETA, int S, int G>
static void add_split(std::vector<SplitInstance>& v) {
    using Geo = SplitGeom<SAT, BM, BN, BK, PM, PN, BETA, S, G>;
    v.push_back(SplitInstance{PREC, BM, BN, BK, S, PM, PN, Geo::NT, Geo::LDS,
                              &launch_split<SAT, Geo>, 1, 0});
    v.push_back(SplitInstance{PREC, BM, BN, BK, S, PM, PN, Geo::NT, Geo::LDS,
                              &launch_split_pipe<SAT, Geo>, 2, 0});
}

template <class SAT, int PREC, int BM, int BN, int BK, int PM, int PN, int BETA, int S, int G,
          int B1>
static void add_split_resident(std::vector<SplitInstance>& v) {
    using Geo = SplitResGeom<SAT, BM, BN, BK, PM, PN, BETA, S, G, B1>;
    v.push_back(SplitInstance{PREC, BM, BN, BK, S, PM, PN, Geo::NT, Geo::LDS,
                              &launch_split_resident<SAT, Geo>, 1, 1});
}

static const std::vector<SplitInstance>& split_registry() {
    static const std::vector<SplitInstance> v = [] {
        std::vector<SplitInstance> r;
        constexpr int F16 = KAMI_PREC_F64_SPLIT_F16, BF = KAMI_PREC_F64_SPLIT_BF16;
        // FP16 digits: 9 bits; default 6 slices (levels <= 7, 21 products)
        add_split<AtomF16x32, F16, 128, 128, 128, 4, 2, 9, 6, 7>(r);
        add_split<AtomF16x32, F16, 128, 128, 128, 4, 2, 9, 7, 8>(r);
        add_split<AtomF16x32, F16, 64, 64, 64, 2, 2, 9, 6, 7>(r);
        add_split<AtomF16x32, F16, 64, 64, 64, 2, 2, 9, 7, 8>(r);
        // BF16 digits: 8 bits; default 7 slices (levels <= 8, 28 products)
        add_split<AtomBF16x32, BF, 128, 128, 128, 4, 2, 8, 7, 8>(r);
        add_split<AtomBF16x32, BF, 64, 64, 64, 2, 2, 8, 7, 8>(r);
        // resident-repeat forms: the whole 64^3 problem's digit planes in LDS, 8 waves;
        // FP16 with an 8-bit first digit, so that every level sum runs over all of K
        add_split_resident<AtomF16x32, F16, 64, 64, 64, 4, 2, 9, 6, 7, 8>(r);
        add_split_resident<AtomBF16x32, BF, 64, 64, 64, 4, 2, 8, 7, 8, 8>(r);
        add_split_resident<AtomF16x32, F16, 32, 32, 32, 2, 2, 9, 6, 7, 8>(r);
        add_split_resident<AtomF16x32, F16, 64, 64, 64, 4, 2, 9, 7, 8, 8>(r);
        add_split_resident<AtomF16x32, F16, 32, 32, 32, 2, 2, 9, 7, 8, 8>(r);
        add_split_resident<AtomBF16x32, BF, 32, 32, 32, 2, 2, 8, 7, 8, 8>(r);
        add_split_i8_instances(r);
        return r;
    }();
    return v;
}

// Default digit count per precision: FP16 7 (9-bit digits, 63 bits), BF16 7 (8-bit
// digits), int8 8 (62-bit fixed point).
static int default_slices(int prec) {
    return prec == KAMI_PREC_F64_SPLIT_I8 ? 8 : 7;
}

// FP16/BF16: pipeline 0 selects the persistent form for iters = 1 and the resident form
// (one workgroup per problem) for iters > 1.  Int8 has one persistent form for any iters.
static const SplitInstance* resolve_split(const kami_gemm_desc* d) {
    const int want = d->slices ? d->slices : default_slices(d->precision);
    const int rep = d->iters > 1 ? 1 : 0;
    const int pipe = d->pipeline ? d->pipeline : (rep ? 1 : 2);
    for (const auto& in : split_registry()) {
        if (in.prec != d->precision || in.m != d->m || in.n != d->n || in.k != d->k ||
            in.slices != want)
            continue;
        if (in.resident == 2) {
            if (d->pipeline == 0 || d->pipeline == in.pipe) return &in;
        } else if (in.pipe == pipe && in.resident == rep) {
            return &in;
        }
    }
    return nullptr;
}

kami_status split_query(const kami_gemm_desc* d, kami_gemm_config* out) {
    if (d->b_layout != KAMI_B_NK || d->partition == KAMI_PART_3D) return KAMI_ENOTSUP;
    const bool i8 = d->precision == KAMI_PREC_F64_SPLIT_I8;
    // int8: each wave owns whole rows of C (1D); FP16/BF16: a 2D wave grid
    if (d->partition == (i8 ? KAMI_PART_2D : KAMI_PART_1D)) return KAMI_ENOTSUP;
    const SplitInstance* in = resolve_split(d);
    if (!in) return KAMI_ENOTSUP;
    out->partition = i8 ? KAMI_PART_1D : KAMI_PART_2D;
    out->ranks = in->pm;
    out->waves_m = in->pm;
    out->waves_n = in->pn;
    out->waves_k = 1;
    out->threads = in->threads;
    out->k_stage = (in->resident || i8) ? in->k : 32;
    out->lds_bytes = in->lds;
    out->slices = in->slices;
    out->instance = (int32_t)(in - split_registry().data());
    out->pipeline = in->pipe;
    out->prefetch = in->resident == 1 ? 0 : 1;
    return KAMI_OK;
}

kami_status split_launch(const kami_gemm_desc* d, const BlockArgs& b, int64_t count,
                         hipStream_t stream) {
    kami_gemm_config cfg;
    kami_status st = split_query(d, &cfg);
    if (st != KAMI_OK) return st;
    const SplitInstance* in = resolve_split(d);
    // one workgroup per problem (pipe 1, resident): work-items must stay below 2^32
    if (in->pipe != 2 && (uint64_t)count * (uint64_t)in->threads >= (1ull << 32))
        return KAMI_ENOTSUP;
    if (d->precision == KAMI_PREC_F64_SPLIT_I8) {
        // the int8 kernel addresses a problem's rows with 32-bit byte offsets
        const int64_t lim = (int64_t)1 << 28;  // elements
        if ((int64_t)(d->m - 1) * b.lda + d->k >= lim || (int64_t)(d->n - 1) * b.ldb + d->k >= lim ||
            (int64_t)(d->m - 1) * b.ldc + d->n >= lim)
            return KAMI_ENOTSUP;
    }
    st = device_check();
    if (st != KAMI_OK) return st;
    SplitArgs a;
    a.A = reinterpret_cast<const double*>(b.A);
    a.B = reinterpret_cast<const double*>(b.B);
    a.C = reinterpret_cast<double*>(b.C);
    a.sa = b.sa;
    a.sb = b.sb;
    a.sc = b.sc;
    a.lda = b.lda;
    a.ldb = b.ldb;
    a.ldc = b.ldc;
    a.alpha = b.alpha;
    a.beta = b.beta;
    a.nch = d->k / 32;
    a.tiles_m = 1;
    a.tiles_n = 1;
    a.iters = b.iters;
    hipError_t e = in->launch(a, count, stream);
    if (e != hipSuccess) return hip_fail(e);
    return KAMI_OK;
}

// Any M, N that are multiples of the tile edge (128, else 64) and K a multiple of 32:
// every problem is cut into tiles of C, each a work unit of the persistent kernel.
kami_status split_gemm_tiled(const kami_gemm_desc* d, const BlockArgs& b, int64_t count,
                             hipStream_t stream) {
    // the int8 form exists for whole block problems only (exponents over the full K)
    if (!is_split(d->precision) || d->precision == KAMI_PREC_F64_SPLIT_I8) return KAMI_ENOTSUP;
    if (d->b_layout != KAMI_B_NK || d->iters != 1) return KAMI_ENOTSUP;
    if (d->partition != KAMI_PART_AUTO && d->partition != KAMI_PART_2D) return KAMI_ENOTSUP;
    if (d->k % 32 != 0) return KAMI_ENOTSUP;
    const int tile = (d->m % 128 == 0 && d->n % 128 == 0) ? 128
                     : (d->m % 64 == 0 && d->n % 64 == 0) ? 64 : 0;
    if (!tile) return KAMI_ENOTSUP;
    kami_gemm_desc t = *d;
    t.m = t.n = t.k = tile;
    t.pipeline = 2;
    const SplitInstance* in = resolve_split(&t);
    if (!in) return KAMI_ENOTSUP;
    const int64_t tiles = (int64_t)(d->m / tile) * (d->n / tile);
    if (count > INT64_MAX / tiles) return KAMI_EINVAL;
    SplitArgs a;
    a.A = reinterpret_cast<const double*>(b.A);
    a.B = reinterpret_cast<const double*>(b.B);
    a.C = reinterpret_cast<double*>(b.C);
    a.sa = b.sa;
    a.sb = b.sb;
    a.sc = b.sc;
    a.lda = b.lda;
    a.ldb = b.ldb;
    a.ldc = b.ldc;
    a.alpha = b.alpha;
    a.beta = b.beta;
    a.nch = d->k / 32;
    a.tiles_m = d->m / tile;
    a.tiles_n = d->n / tile;
    a.iters = 1;
    kami_status st = device_check();
    if (st != KAMI_OK) return st;
    hipError_t e = in->launch(a, count * tiles, stream);
    if (e != hipSuccess) return hip_fail(e);
    return KAMI_OK;
}

}  // namespace kami
