// Mixed-precision blocked Cholesky (no counterpart in the reference; its
// nearest relative is the whole-program FP32 build, CMakeLists.txt:7,104-106).
//
// The factor is stored in FP32 (packed layout, 4 bytes per entry). Per step k:
//  - POTRF of the diagonal block and TRSM of the panel run in FP64 arithmetic
//    on the FP32-stored values (the FP64 kernels, templated on storage);
//  - the trailing SYRK/GEMM update C -= L_Ik L_Jk^T runs on MFMA with FP32
//    accumulation, in one of two variants:
//      HS_MIXED_FP32:   v_mfma_f32_16x16x4f32 on the FP32 values;
//      HS_MIXED_BF16X3: every FP32 operand x is split into bf16 hi = rn(x) and
//                       lo = rn(x - hi); each product uses three
//                       v_mfma_f32_16x16x32_bf16: hi*hi' + hi*lo' + lo*hi'
//                       (the dropped lo*lo' and the split residuals are
//                       ~2^-17 relative).
//      HS_MIXED_F16X3:  the same three products on v_mfma_f32_16x16x32_f16
//                       with an fp16 split of s*x, s a power of two that puts
//                       the largest panel entry just under 2^15: fp16 keeps
//                       11 significand bits to bf16's 8, so hi + lo holds
//                       ~22 bits and the update is near FP32 accuracy at the
//                       BF16 MFMA rate (the FP16 hi+lo construction of Ootomo
//                       and Yokota, "Recovering single precision accuracy from
//                       Tensor Cores while surpassing the FP32 theoretical
//                       peak performance", IJHPCA 2022, with one power-of-two
//                       scale of the whole factor in place of their scaled lo
//                       term). The accumulator is scaled back by 1/s^2, exact.
// The steps follow the same one-panel look-ahead schedule as the FP64 path.
#include <algorithm>
#include <cstdlib>

#include "hs_factor.h"
#include "hs_internal.h"


namespace {

constexpr int THREADS = 256;
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

// K (FP64 packed) -> FP32 packed copy, + shift on the global diagonal.
// vec: four elements per thread and step (two 16-byte loads, one 16-byte
// store) for 16-byte aligned buffers; the elements past the last multiple of
// 4, and unaligned buffers, one at a time.
typedef double dbl2 __attribute__((ext_vector_type(2)));
__global__ void to_f32_kernel(const double* __restrict__ K, float* __restrict__ L, i64 total, int vec) {
    const i64 gid = (i64)blockIdx.x * blockDim.x + threadIdx.x, stride = (i64)gridDim.x * blockDim.x;
    const i64 q = vec ? total / 4 : 0;
    for (i64 e = gid; e < q; e += stride) {
        const dbl2 a = __builtin_nontemporal_load(reinterpret_cast<const dbl2*>(K) + 2 * e);
        const dbl2 c = __builtin_nontemporal_load(reinterpret_cast<const dbl2*>(K) + 2 * e + 1);
        reinterpret_cast<f4*>(L)[e] = f4{(float)a.x, (float)a.y, (float)c.x, (float)c.y};
    }
    for (i64 e = 4 * q + gid; e < total; e += stride) L[e] = (float)K[e];
}

__global__ void diag_shift_kernel(const double* __restrict__ K, float* __restrict__ L, int b, int n, i64 N,
                                  double shift) {
    const int j = blockIdx.x;
    for (int i = threadIdx.x; i < b; i += blockDim.x) {
        if ((i64)j * b + i >= N) continue;
        const i64 e = hs_block_id(j, j, n) * (i64)b * b + (i64)i * b + i;
        L[e] = (float)(K[e] + shift);
    }
}

// ---- FP32 MFMA trailing update -------------------------------------------
// C -= acc for one work-group tile (f32 MFMA layout: register v of lane l is
// row 4(l/16) + v, column l%16). The tile is read in batches of one 16-row
// group with the next group's loads issued before the current group's
// stores: the compiler cannot prove the stores do not alias later loads and
// would otherwise wait for every load on its own. On the diagonal tile of a
// diagonal block only the lower triangle is written.
template <int MI, int MJ>
__device__ inline void epilogue_f32(float* C, int b, int wm0, int wn0, int lr, int lq, bool dtile,
                                    const f4 (&acc)[MI][MJ]) {
    float* pc = C + (i64)(wm0 + 4 * lq) * b + wn0 + lr;
    float cv[MJ][4], cn[MJ][4];
#pragma unroll
    for (int j = 0; j < MJ; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) cv[j][v] = pc[(i64)v * b + j * 16];
#pragma unroll
    for (int i = 0; i < MI; ++i) {
        if (i + 1 < MI) {
#pragma unroll
            for (int j = 0; j < MJ; ++j)
#pragma unroll
                for (int v = 0; v < 4; ++v) cn[j][v] = pc[(i64)((i + 1) * 16 + v) * b + j * 16];
        }
#pragma unroll
        for (int j = 0; j < MJ; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = wm0 + i * 16 + 4 * lq + v, col = wn0 + j * 16 + lr;
                if (!dtile || row >= col) pc[(i64)(i * 16 + v) * b + j * 16] = cv[j][v] - acc[i][j][v];
            }
        if (i + 1 < MI) {
#pragma unroll
            for (int j = 0; j < MJ; ++j)
#pragma unroll
                for (int v = 0; v < 4; ++v) cv[j][v] = cn[j][v];
        }
    }
}

// TILE x TILE output tile per work-group, 2x2 waves, 16x16x4 f32 MFMA tiles
// (register v of lane l: row 4(l/16)+v, column l%16), KT = 16 deep LDS
// stages, double buffered, operands stored k-major.
template <int TILE, int KIND>
__global__ __launch_bounds__(THREADS) void trailing_mfma_f32(float* __restrict__ A, int b, int n, PairMap pm) {
    constexpr int KT = 16;
    constexpr int WT = TILE / 2, MT = WT / 16;
    constexpr int LDP = TILE + 4;
    constexpr int STAGE = KT * LDP;
    __shared__ float As[2 * STAGE];
    __shared__ float Bs[2 * STAGE];
    const int tpb = (b / TILE) * (b / TILE);
    int I, J, tile;
    if (!tile_of(pm, tpb, &I, &J, &tile)) return;
    const int tm = tile / (b / TILE), tn = tile % (b / TILE);
    const bool diag = (I == J);
    if (diag && tm < tn) return;
    const i64 bb = (i64)b * b;
    const float* La = A + hs_block_id(I, pm.k, n) * bb + (i64)tm * TILE * b;
    const float* Lb = A + hs_block_id(J, pm.k, n) * bb + (i64)tn * TILE * b;
    float* C = A + hs_block_id(I, J, n) * bb + (i64)tm * TILE * b + (i64)tn * TILE;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm0 = (wave >> 1) * WT, wn0 = (wave & 1) * WT;
    constexpr int LPT = TILE * KT / THREADS;  // 8 floats per thread (TILE 128)
    constexpr int TPR = KT / LPT;
    const int lrow = tid / TPR, lk = (tid % TPR) * LPT;
    float4 ra[LPT / 4], rb[LPT / 4];
    const float* pa = La + (i64)lrow * b + lk;
    const float* pb = Lb + (i64)lrow * b + lk;
    f4 acc[MT][MT];
#pragma unroll
    for (int i = 0; i < MT; ++i)
#pragma unroll
        for (int j = 0; j < MT; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
    // np panels: in column-major block order, block (I, k+p) lies
    // p (n-k-1) - p (p-1)/2 blocks after block (I, k), the same for every row
    const int nk = b / KT, nkt = nk * pm.np;
    auto poff = [&](int kt) -> i64 {
        const i64 pp = kt / nk;
        return (pp * (n - pm.k - 1) - pp * (pp - 1) / 2) * bb + (i64)(kt % nk) * KT;
    };
    auto gload = [&](int kt) {
#pragma unroll
        for (int v = 0; v < LPT / 4; ++v) {
            ra[v] = *reinterpret_cast<const float4*>(pa + poff(kt) + 4 * v);
            rb[v] = *reinterpret_cast<const float4*>(pb + poff(kt) + 4 * v);
        }
    };
    auto sstore = [&](int buf) {
        float* as = As + buf * STAGE;
        float* bs = Bs + buf * STAGE;
#pragma unroll
        for (int v = 0; v < LPT / 4; ++v) {
            as[(lk + 4 * v) * LDP + lrow] = ra[v].x;
            as[(lk + 4 * v + 1) * LDP + lrow] = ra[v].y;
            as[(lk + 4 * v + 2) * LDP + lrow] = ra[v].z;
            as[(lk + 4 * v + 3) * LDP + lrow] = ra[v].w;
            bs[(lk + 4 * v) * LDP + lrow] = rb[v].x;
            bs[(lk + 4 * v + 1) * LDP + lrow] = rb[v].y;
            bs[(lk + 4 * v + 2) * LDP + lrow] = rb[v].z;
            bs[(lk + 4 * v + 3) * LDP + lrow] = rb[v].w;
        }
    };
    gload(0);
    sstore(0);
    if (nkt > 1) gload(1);
    __syncthreads();
    const int lr = lane & 15, lq = lane >> 4;
    for (int kt = 0; kt < nkt; ++kt) {
        const float* as = As + (kt & 1) * STAGE;
        const float* bs = Bs + (kt & 1) * STAGE;
#pragma unroll
        for (int ks = 0; ks < KT; ks += 4) {
            float av[MT], bv[MT];
#pragma unroll
            for (int i = 0; i < MT; ++i) av[i] = as[(ks + lq) * LDP + wm0 + i * 16 + lr];
#pragma unroll
            for (int j = 0; j < MT; ++j) bv[j] = bs[(ks + lq) * LDP + wn0 + j * 16 + lr];
            // the other buffer was last read in stage kt-1 (see trailing_mfma_f64)
            if (ks == 0 && kt + 1 < nkt) {
                sstore((kt + 1) & 1);
                if (kt + 2 < nkt) gload(kt + 2);
            }
#pragma unroll
            for (int i = 0; i < MT; ++i)
#pragma unroll
                for (int j = 0; j < MT; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[i], bv[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
    epilogue_f32<MT, MT>(C, b, wm0, wn0, lr, lq, diag && tm == tn, acc);
}

// ---- hi+lo split trailing updates (BF16 and scaled FP16) -------------------
__device__ inline unsigned short bf16_rn(float x) {
    const unsigned u = __float_as_uint(x);
    return (unsigned short)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ inline float bf16_to_f(unsigned short h) { return __uint_as_float(((unsigned)h) << 16); }

// The 16-bit format of a split: rn rounds an FP32 value to it (nearest,
// ties to even), to_f widens it back exactly, mfma is the 16x16x32 MFMA
// with FP32 accumulation. An fp16 split works on s*x (Hilo::scaled): sc[0]
// holds s and sc[1] 1/s^2, both powers of two (hilo_scale_finish_kernel).
enum class H16 { BF16, F16 };
template <H16 F>
struct Hilo;
template <>
struct Hilo<H16::BF16> {
    static constexpr bool scaled = false;
    typedef __bf16 frag __attribute__((ext_vector_type(8)));
    static __device__ inline unsigned short rn(float x) { return bf16_rn(x); }
    static __device__ inline float to_f(unsigned short h) { return bf16_to_f(h); }
    static __device__ inline f4 mfma(frag a, frag b, f4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
};
template <>
struct Hilo<H16::F16> {
    static constexpr bool scaled = true;
    typedef _Float16 frag __attribute__((ext_vector_type(8)));
    static __device__ inline unsigned short rn(float x) { return __builtin_bit_cast(unsigned short, (_Float16)x); }
    static __device__ inline float to_f(unsigned short h) { return (float)__builtin_bit_cast(_Float16, h); }
    static __device__ inline f4 mfma(frag a, frag b, f4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
};

// hi = rn(s x), lo = rn(s x - hi): s x and the difference are exact in FP32
template <H16 F>
__device__ inline void hilo_split(float x, float s, unsigned short* hi, unsigned short* lo) {
    const float xs = Hilo<F>::scaled ? x * s : x;
    const unsigned short h = Hilo<F>::rn(xs);
    *hi = h;
    *lo = Hilo<F>::rn(xs - Hilo<F>::to_f(h));
}

// the accumulated products of scaled operands back to the matrix's scale
template <H16 F, int MI, int MJ>
__device__ inline void hilo_unscale(f4 (&acc)[MI][MJ], const float* sc) {
    if constexpr (Hilo<F>::scaled) {
        const float r = sc[1];
#pragma unroll
        for (int i = 0; i < MI; ++i)
#pragma unroll
            for (int j = 0; j < MJ; ++j) acc[i][j] *= r;
    }
}

// TILE x TILE tile per work-group, 2x2 waves of (TILE/2)^2, KT = 32 (one
// 16-bit MFMA depth) per LDS stage, double buffered. LDS keeps hi and lo
// planes row-major [row][k] with an 80-byte row pitch, so the 16-byte
// fragment reads of 16 lanes (8 consecutive k of 16 rows) hit distinct banks.
// Fragment of lane l: row l%16, k = 8(l/16)..+7; output as the f32 MFMA.
//
// PS (pre-split): the operand blocks come from hi / lo planes that
// split_panels_kernel wrote once for the whole launch (Wh, Wl: np panels x
// m block rows from rs, b x b 16-bit values each), so the stage fill is a
// copy; otherwise every work-group splits the FP32 values it loads, which
// costs more VALU issue than the MFMAs of the stage.
template <int TILE, bool PS, H16 F>
__device__ __forceinline__ void trailing_x3_body(float* __restrict__ A, int b, int n, const PairMap& pm,
                                                 const unsigned short* __restrict__ Wh,
                                                 const unsigned short* __restrict__ Wl, int rs, int m,
                                                 const float* __restrict__ sc) {
    typedef typename Hilo<F>::frag bf8;
    constexpr int KT = 32;
    constexpr int WT = TILE / 2, MT = WT / 16;
    constexpr int P = 40;                 // row pitch in 16-bit elements (80 bytes)
    constexpr int PLANE = TILE * P;       // one hi or lo plane of one operand
    constexpr int STAGE = 4 * PLANE;      // A hi, A lo, B hi, B lo
    __shared__ unsigned short lds[2 * STAGE];
    const int tpb = (b / TILE) * (b / TILE);
    int I, J, tile;
    if (!tile_of(pm, tpb, &I, &J, &tile)) return;
    const int tm = tile / (b / TILE), tn = tile % (b / TILE);
    const bool diag = (I == J);
    if (diag && tm < tn) return;
    const i64 bb = (i64)b * b;
    const float* La = A + hs_block_id(I, pm.k, n) * bb + (i64)tm * TILE * b;
    const float* Lb = A + hs_block_id(J, pm.k, n) * bb + (i64)tn * TILE * b;
    float* C = A + hs_block_id(I, J, n) * bb + (i64)tm * TILE * b + (i64)tn * TILE;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm0 = (wave >> 1) * WT, wn0 = (wave & 1) * WT;
    constexpr int LPT = TILE * KT / THREADS;  // 16 floats per thread (TILE 128)
    constexpr int TPR = KT / LPT;             // 2 threads per row
    const int lrow = tid / TPR, lk = (tid % TPR) * LPT;
    float4 ra[PS ? 1 : LPT / 4], rb[PS ? 1 : LPT / 4];
    u32x4 rah[PS ? LPT / 8 : 1], ral[PS ? LPT / 8 : 1], rbh[PS ? LPT / 8 : 1], rbl[PS ? LPT / 8 : 1];
    const float* pa = La + (i64)lrow * b + lk;
    const float* pb = Lb + (i64)lrow * b + lk;
    // plane offsets of this thread's rows in block rows I and J of panel 0
    const i64 wa = PS ? ((i64)(I - rs) * bb + (i64)(tm * TILE + lrow) * b + lk) : 0;
    const i64 wb = PS ? ((i64)(J - rs) * bb + (i64)(tn * TILE + lrow) * b + lk) : 0;
    f4 acc[MT][MT];
#pragma unroll
    for (int i = 0; i < MT; ++i)
#pragma unroll
        for (int j = 0; j < MT; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
    // np panels: in column-major block order, block (I, k+p) lies
    // p (n-k-1) - p (p-1)/2 blocks after block (I, k), the same for every row
    const int nk = b / KT, nkt = nk * pm.np;
    auto poff = [&](int kt) -> i64 {
        const i64 pp = kt / nk;
        return (pp * (n - pm.k - 1) - pp * (pp - 1) / 2) * bb + (i64)(kt % nk) * KT;
    };
    auto gload = [&](int kt) {
        if constexpr (PS) {
            const i64 o = (i64)(kt / nk) * m * bb + (i64)(kt % nk) * KT;
#pragma unroll
            for (int v = 0; v < LPT / 8; ++v) {
                rah[v] = *reinterpret_cast<const u32x4*>(Wh + wa + o + 8 * v);
                ral[v] = *reinterpret_cast<const u32x4*>(Wl + wa + o + 8 * v);
                rbh[v] = *reinterpret_cast<const u32x4*>(Wh + wb + o + 8 * v);
                rbl[v] = *reinterpret_cast<const u32x4*>(Wl + wb + o + 8 * v);
            }
        } else {
#pragma unroll
            for (int v = 0; v < LPT / 4; ++v) {
                ra[v] = *reinterpret_cast<const float4*>(pa + poff(kt) + 4 * v);
                rb[v] = *reinterpret_cast<const float4*>(pb + poff(kt) + 4 * v);
            }
        }
    };
    const float scale = (!PS && Hilo<F>::scaled) ? sc[0] : 1.0f;
    auto split_store = [&](const float4* r, unsigned short* hiPlane, unsigned short* loPlane) {
        unsigned short hi[LPT], lo[LPT];
#pragma unroll
        for (int v = 0; v < LPT / 4; ++v) {
            const float x[4] = {r[v].x, r[v].y, r[v].z, r[v].w};
#pragma unroll
            for (int c = 0; c < 4; ++c) hilo_split<F>(x[c], scale, &hi[4 * v + c], &lo[4 * v + c]);
        }
        uint4* dh = reinterpret_cast<uint4*>(hiPlane + lrow * P + lk);
        uint4* dl = reinterpret_cast<uint4*>(loPlane + lrow * P + lk);
#pragma unroll
        for (int w = 0; w < LPT / 8; ++w) {
            uint4 vh, vl;
            vh.x = hi[8 * w] | ((unsigned)hi[8 * w + 1] << 16);
            vh.y = hi[8 * w + 2] | ((unsigned)hi[8 * w + 3] << 16);
            vh.z = hi[8 * w + 4] | ((unsigned)hi[8 * w + 5] << 16);
            vh.w = hi[8 * w + 6] | ((unsigned)hi[8 * w + 7] << 16);
            vl.x = lo[8 * w] | ((unsigned)lo[8 * w + 1] << 16);
            vl.y = lo[8 * w + 2] | ((unsigned)lo[8 * w + 3] << 16);
            vl.z = lo[8 * w + 4] | ((unsigned)lo[8 * w + 5] << 16);
            vl.w = lo[8 * w + 6] | ((unsigned)lo[8 * w + 7] << 16);
            dh[w] = vh;
            dl[w] = vl;
        }
    };
    auto sstore = [&](int buf) {
        unsigned short* s = lds + buf * STAGE;
        if constexpr (PS) {
            u32x4* d = reinterpret_cast<u32x4*>(s + lrow * P + lk);
#pragma unroll
            for (int v = 0; v < LPT / 8; ++v) {
                d[v] = rah[v];
                d[PLANE / 8 + v] = ral[v];
                d[2 * PLANE / 8 + v] = rbh[v];
                d[3 * PLANE / 8 + v] = rbl[v];
            }
        } else {
            split_store(ra, s, s + PLANE);
            split_store(rb, s + 2 * PLANE, s + 3 * PLANE);
        }
    };
    gload(0);
    sstore(0);
    if (nkt > 1) gload(1);
    __syncthreads();
    const int lr = lane & 15, lq = lane >> 4;
    for (int kt = 0; kt < nkt; ++kt) {
        const unsigned short* s = lds + (kt & 1) * STAGE;
        bf8 ah[MT], al[MT], bh[MT], bl[MT];
#pragma unroll
        for (int i = 0; i < MT; ++i) {
            const int off = (wm0 + i * 16 + lr) * P + 8 * lq;
            ah[i] = *reinterpret_cast<const bf8*>(s + off);
            al[i] = *reinterpret_cast<const bf8*>(s + PLANE + off);
        }
#pragma unroll
        for (int j = 0; j < MT; ++j) {
            const int off = (wn0 + j * 16 + lr) * P + 8 * lq;
            bh[j] = *reinterpret_cast<const bf8*>(s + 2 * PLANE + off);
            bl[j] = *reinterpret_cast<const bf8*>(s + 3 * PLANE + off);
        }
        // the other buffer was last read in stage kt-1: fill it while this
        // stage's MFMAs run
        if (kt + 1 < nkt) {
            sstore((kt + 1) & 1);
            if (kt + 2 < nkt) gload(kt + 2);
        }
#pragma unroll
        for (int i = 0; i < MT; ++i)
#pragma unroll
            for (int j = 0; j < MT; ++j) {
                acc[i][j] = Hilo<F>::mfma(al[i], bh[j], acc[i][j]);
                acc[i][j] = Hilo<F>::mfma(ah[i], bl[j], acc[i][j]);
                acc[i][j] = Hilo<F>::mfma(ah[i], bh[j], acc[i][j]);
            }
        __syncthreads();
    }
    hilo_unscale<F>(acc, sc);
    epilogue_f32<MT, MT>(C, b, wm0, wn0, lr, lq, diag && tm == tn, acc);
}

// KIND only separates the look-ahead column launches (KIND 1, modes 5 and 6)
// from the rest in profiles; the code is the same.
template <int TILE, int KIND, bool PS>
__global__ __launch_bounds__(THREADS) void trailing_mfma_bf16x3(float* __restrict__ A, int b, int n, PairMap pm,
                                                                 const unsigned short* __restrict__ Wh,
                                                                 const unsigned short* __restrict__ Wl, int rs,
                                                                 int m) {
    trailing_x3_body<TILE, PS, H16::BF16>(A, b, n, pm, Wh, Wl, rs, m, nullptr);
}
template <int TILE, int KIND, bool PS>
__global__ __launch_bounds__(THREADS) void trailing_mfma_f16x3(float* __restrict__ A, int b, int n, PairMap pm,
                                                                const unsigned short* __restrict__ Wh,
                                                                const unsigned short* __restrict__ Wl, int rs, int m,
                                                                const float* __restrict__ sc) {
    trailing_x3_body<TILE, PS, H16::F16>(A, b, n, pm, Wh, Wl, rs, m, sc);
}

// Rest update from pre-split planes in 256 x 256 tiles (b % 256 == 0): 512
// threads as 4 x 2 waves of 64 x 128, KT = 32 deep stages of the four planes
// (A hi, A lo, B hi, B lo; 16 KB each), double-buffered in 128 KB of dynamic
// LDS. Half the operand traffic per MFMA of the 128 tile.
//  - Stages are filled by LDS DMA (global_load_lds_dwordx4: each lane's 16
//    bytes land at the wave's LDS base + 16 * lane), so no VGPR staging and
//    no LDS stores. Planes are unpadded, 64 bytes per row, with the 16-byte
//    k-chunk q of row r in slot q ^ ((r >> 2) & 3): the fragment reads of 16
//    lanes (rows r .. r+15, one chunk) then cover all 64 banks once. Wave w
//    fills plane w / 2, rows (w & 1) * 128 .. +127, 16 rows per DMA, lane L
//    taking row L / 4 and chunk (L & 3) ^ (L >> 4), which is slot L & 3.
//  - A stage runs as two halves of 64 output columns. The second half's B
//    fragments are read while the first half's MFMAs run; then one barrier
//    (this wave's DMA of the next stage has landed, every wave's reads of the
//    stage are done), the DMA of stage kt+2 into the buffer just consumed,
//    and the next stage's A and first-half B fragments are read while the
//    second half's MFMAs run. The barrier carries no memory fence: a fence
//    would wait for the DMA in flight, which the explicit counters cover.
constexpr int T2 = 256, T2THREADS = 512;
constexpr int T2PLANE = T2 * 64;                    // bytes per plane and stage
constexpr int T2STAGE = 4 * T2PLANE;
constexpr size_t T2LDS = 2 * (size_t)T2STAGE;
constexpr unsigned WAIT_VM0_LGKM0 = 0x0070;         // s_waitcnt vmcnt(0) lgkmcnt(0)
constexpr unsigned WAIT_VM0 = 0x0F70;               // s_waitcnt vmcnt(0)

template <H16 F>
__device__ __forceinline__ void trailing_x3_ps256_body(float* __restrict__ A, int b, int n, const PairMap& pm,
                                                       const unsigned short* __restrict__ Wh,
                                                       const unsigned short* __restrict__ Wl, int rs, int m,
                                                       const float* __restrict__ sc) {
    typedef typename Hilo<F>::frag bf8;
    constexpr int KT = 32;
    extern __shared__ __attribute__((aligned(16))) unsigned char lds2[];
    const int tpe = b / T2, tpb = tpe * tpe;
    int I, J, tile;
    if (!tile_of(pm, tpb, &I, &J, &tile)) return;
    const int tm = tile / tpe, tn = tile % tpe;
    const bool diag = (I == J);
    if (diag && tm < tn) return;
    const i64 bb = (i64)b * b;
    float* C = A + hs_block_id(I, J, n) * bb + (i64)tm * T2 * b + (i64)tn * T2;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm0 = (wave >> 1) * 64, wn0 = (wave & 1) * 128;
    const int plane = wave >> 1;
    const unsigned short* src = (plane & 1) ? Wl : Wh;
    const int rsel = plane < 2 ? I : J, tsel = plane < 2 ? tm : tn;
    const int rr = (wave & 1) * 128 + (lane >> 2);
    const int q = (lane & 3) ^ (lane >> 4);
    const unsigned short* gp = src + (i64)(rsel - rs) * bb + (i64)(tsel * T2 + rr) * b + q * 8;
    unsigned char* lb = lds2 + plane * T2PLANE + (wave & 1) * 8 * 1024;
    const int nk = b / KT, nkt = nk * pm.np;
    auto dma = [&](int kt, int buf) {
        const i64 o = (i64)(kt / nk) * m * bb + (i64)(kt % nk) * KT;
#pragma unroll
        for (int j = 0; j < 8; ++j)
            __builtin_amdgcn_global_load_lds(gp + o + (i64)j * 16 * b, lb + buf * T2STAGE + j * 1024, 16, 0, 0);
    };
    const int lr = lane & 15, lq = lane >> 4;
    const int sw = (lq ^ (lr >> 2)) * 16;
    bf8 ah[4], al[4], b0h[4], b0l[4], b1h[4], b1l[4];
    auto readA = [&](const unsigned char* s) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int off = (wm0 + i * 16 + lr) * 64 + sw;
            ah[i] = *reinterpret_cast<const bf8*>(s + off);
            al[i] = *reinterpret_cast<const bf8*>(s + T2PLANE + off);
        }
    };
    auto readB = [&](const unsigned char* s, int h, bf8 (&bh)[4], bf8 (&bl)[4]) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int off = (wn0 + (4 * h + j) * 16 + lr) * 64 + sw;
            bh[j] = *reinterpret_cast<const bf8*>(s + 2 * T2PLANE + off);
            bl[j] = *reinterpret_cast<const bf8*>(s + 3 * T2PLANE + off);
        }
    };
    f4 acc[4][8];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
    auto half = [&](int h, const bf8 (&bh)[4], const bf8 (&bl)[4]) {
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                f4& c = acc[i][4 * h + j];
                c = Hilo<F>::mfma(al[i], bh[j], c);
                c = Hilo<F>::mfma(ah[i], bl[j], c);
                c = Hilo<F>::mfma(ah[i], bh[j], c);
            }
    };
    dma(0, 0);
    __builtin_amdgcn_s_waitcnt(WAIT_VM0);
    __syncthreads();
    if (nkt > 1) dma(1, 1);
    readA(lds2);
    readB(lds2, 0, b0h, b0l);
    for (int kt = 0; kt < nkt; ++kt) {
        const unsigned char* s = lds2 + (kt & 1) * T2STAGE;
        readB(s, 1, b1h, b1l);
        half(0, b0h, b0l);
        // 48 MFMAs, the 8 reads spread over the first 16
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);
            __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
        }
        __builtin_amdgcn_sched_group_barrier(0x008, 32, 0);
        __builtin_amdgcn_s_waitcnt(WAIT_VM0_LGKM0);
        asm volatile("" ::: "memory");
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
        // unconditional, so that the stage stays one basic block for the
        // schedule above and below: past the last stage the DMA refetches
        // stage 0 into the buffer just consumed, and the reads take the other
        // buffer's contents, which are never used
        dma(kt + 2 < nkt ? kt + 2 : 0, kt & 1);
        half(1, b1h, b1l);
        {
            const unsigned char* s1 = lds2 + ((kt + 1) & 1) * T2STAGE;
            readB(s1, 0, b0h, b0l);
            readA(s1);
        }
        // 48 MFMAs with the 16 reads spread over the first 32
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            __builtin_amdgcn_sched_group_barrier(0x008, 2, 1);
            __builtin_amdgcn_sched_group_barrier(0x100, 1, 1);
        }
        __builtin_amdgcn_sched_group_barrier(0x008, 16, 1);
    }
    __builtin_amdgcn_s_waitcnt(WAIT_VM0);  // no DMA into LDS outlives the work-group
    hilo_unscale<F>(acc, sc);
    epilogue_f32<4, 8>(C, b, wm0, wn0, lr, lq, diag && tm == tn, acc);
}

__global__ __launch_bounds__(T2THREADS) void trailing_bf16x3_ps256(float* __restrict__ A, int b, int n, PairMap pm,
                                                                   const unsigned short* __restrict__ Wh,
                                                                   const unsigned short* __restrict__ Wl, int rs,
                                                                   int m) {
    trailing_x3_ps256_body<H16::BF16>(A, b, n, pm, Wh, Wl, rs, m, nullptr);
}
__global__ __launch_bounds__(T2THREADS) void trailing_f16x3_ps256(float* __restrict__ A, int b, int n, PairMap pm,
                                                                  const unsigned short* __restrict__ Wh,
                                                                  const unsigned short* __restrict__ Wl, int rs,
                                                                  int m, const float* __restrict__ sc) {
    trailing_x3_ps256_body<H16::F16>(A, b, n, pm, Wh, Wl, rs, m, sc);
}

// FP32 rest update in 256 x 256 tiles (b % 256 == 0), the FP32 sibling of
// trailing_bf16x3_ps256: 512 threads as 4 x 2 waves of 64 x 128, KT = 16
// deep stages of the two operand planes (A rows of block (I, k+p), B rows of
// block (J, k+p); 16 KB each, 64-byte rows with the same chunk swizzle)
// filled by LDS DMA straight from the FP32 panels, double-buffered in 64 KB.
// A lane's 16-byte fragment read holds four consecutive k; MFMA step s of a
// stage contracts k = 4 g + s over the lane groups g, the same for A and B.
// One work-group fills a CU (two waves per SIMD), so a panel kernel finds a
// free CU whenever one of them finishes.
constexpr int F2PLANE = T2 * 64;                    // bytes per plane and stage
constexpr int F2STAGE = 2 * F2PLANE;
constexpr size_t F2LDS = 2 * (size_t)F2STAGE;

__global__ __launch_bounds__(T2THREADS) void trailing_f32_ps256(float* __restrict__ A, int b, int n, PairMap pm) {
    constexpr int KT = 16;
    extern __shared__ __attribute__((aligned(16))) unsigned char lds3[];
    const int tpe = b / T2, tpb = tpe * tpe;
    int I, J, tile;
    if (!tile_of(pm, tpb, &I, &J, &tile)) return;
    const int tm = tile / tpe, tn = tile % tpe;
    const bool diag = (I == J);
    if (diag && tm < tn) return;
    const i64 bb = (i64)b * b;
    float* C = A + hs_block_id(I, J, n) * bb + (i64)tm * T2 * b + (i64)tn * T2;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm0 = (wave >> 1) * 64, wn0 = (wave & 1) * 128;
    // staging: waves 0-3 fill the A plane, 4-7 the B plane, 64 rows each, 16
    // rows per DMA; lane L takes row L / 4 of the 16 and k-chunk
    // (L & 3) ^ (L >> 4), which lands in slot L & 3
    const int plane = wave >> 2;
    const int rr = (wave & 3) * 64 + (lane >> 2);
    const int q = (lane & 3) ^ (lane >> 4);
    const float* gp = A + hs_block_id(plane ? J : I, pm.k, n) * bb + (i64)((plane ? tn : tm) * T2 + rr) * b + q * 4;
    unsigned char* lb = lds3 + plane * F2PLANE + (wave & 3) * 4 * 1024;
    const int nk = b / KT, nkt = nk * pm.np;
    // np panels: block (I, k+p) lies p (n-k-1) - p (p-1)/2 blocks after block (I, k)
    auto koff = [&](int kt) -> i64 {
        const i64 pp = kt / nk;
        return (pp * (n - pm.k - 1) - pp * (pp - 1) / 2) * bb + (i64)(kt % nk) * KT;
    };
    auto dma = [&](int kt, int buf) {
        const i64 o = koff(kt);
#pragma unroll
        for (int j = 0; j < 4; ++j)
            __builtin_amdgcn_global_load_lds(gp + o + (i64)j * 16 * b, lb + buf * F2STAGE + j * 1024, 16, 0, 0);
    };
    const int lr = lane & 15, lq = lane >> 4;
    const int sw = (lq ^ (lr >> 2)) * 16;
    f4 a[4], b0[4], b1[4];
    auto readA = [&](const unsigned char* st) {
#pragma unroll
        for (int i = 0; i < 4; ++i) a[i] = *reinterpret_cast<const f4*>(st + (wm0 + i * 16 + lr) * 64 + sw);
    };
    auto readB = [&](const unsigned char* st, int h, f4 (&bv)[4]) {
#pragma unroll
        for (int j = 0; j < 4; ++j)
            bv[j] = *reinterpret_cast<const f4*>(st + F2PLANE + (wn0 + (4 * h + j) * 16 + lr) * 64 + sw);
    };
    f4 acc[4][8];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
    auto half = [&](int h, const f4 (&bv)[4]) {
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    acc[i][4 * h + j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i][s], bv[j][s], acc[i][4 * h + j], 0, 0, 0);
    };
    dma(0, 0);
    __builtin_amdgcn_s_waitcnt(WAIT_VM0);
    __syncthreads();
    if (nkt > 1) dma(1, 1);
    readA(lds3);
    readB(lds3, 0, b0);
    for (int kt = 0; kt < nkt; ++kt) {
        const unsigned char* st = lds3 + (kt & 1) * F2STAGE;
        readB(st, 1, b1);
        half(0, b0);
        __builtin_amdgcn_s_waitcnt(WAIT_VM0_LGKM0);
        asm volatile("" ::: "memory");
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
        // unconditional, to keep the stage one basic block: past the last
        // stage the DMA refetches stage 0 into the buffer just consumed and
        // the reads take the other buffer's contents, never used
        dma(kt + 2 < nkt ? kt + 2 : 0, kt & 1);
        half(1, b1);
        {
            const unsigned char* s1 = lds3 + ((kt + 1) & 1) * F2STAGE;
            readB(s1, 0, b0);
            readA(s1);
        }
    }
    __builtin_amdgcn_s_waitcnt(WAIT_VM0);  // no DMA into LDS outlives the work-group
    epilogue_f32<4, 8>(C, b, wm0, wn0, lr, lq, diag && tm == tn, acc);
}

// The hi / lo planes of np panels (block columns k .. k+np-1) for the m
// block rows from rs: W[(p m + i) b^2 + e] for block (rs + i, k + p),
// hilo_split of each value, eight values per thread.
template <H16 F>
__device__ __forceinline__ void split_panels_body(const float* __restrict__ L, int b, int n, int k, int np, int rs,
                                                  int m, unsigned short* __restrict__ Wh,
                                                  unsigned short* __restrict__ Wl, const float* __restrict__ sc) {
    const float scale = Hilo<F>::scaled ? sc[0] : 1.0f;
    const i64 bb = (i64)b * b;
    const i64 total = (i64)np * m * bb;
    for (i64 e = ((i64)blockIdx.x * blockDim.x + threadIdx.x) * 8; e < total; e += (i64)gridDim.x * blockDim.x * 8) {
        const i64 blk = e / bb, off = e % bb;
        const int p = (int)(blk / m), i = (int)(blk % m);
        const float* src = L + hs_block_id(rs + i, k + p, n) * bb + off;
        const float4 x0 = *reinterpret_cast<const float4*>(src);
        const float4 x1 = *reinterpret_cast<const float4*>(src + 4);
        const float x[8] = {x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w};
        unsigned hi[4], lo[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            unsigned short h0, h1, l0, l1;
            hilo_split<F>(x[2 * c], scale, &h0, &l0);
            hilo_split<F>(x[2 * c + 1], scale, &h1, &l1);
            hi[c] = h0 | ((unsigned)h1 << 16);
            lo[c] = l0 | ((unsigned)l1 << 16);
        }
        *reinterpret_cast<uint4*>(Wh + e) = make_uint4(hi[0], hi[1], hi[2], hi[3]);
        *reinterpret_cast<uint4*>(Wl + e) = make_uint4(lo[0], lo[1], lo[2], lo[3]);
    }
}

__global__ __launch_bounds__(256) void split_panels_bf16_kernel(const float* __restrict__ L, int b, int n, int k,
                                                                int np, int rs, int m, unsigned short* __restrict__ Wh,
                                                                unsigned short* __restrict__ Wl) {
    split_panels_body<H16::BF16>(L, b, n, k, np, rs, m, Wh, Wl, nullptr);
}
__global__ __launch_bounds__(256) void split_panels_f16_kernel(const float* __restrict__ L, int b, int n, int k,
                                                               int np, int rs, int m, unsigned short* __restrict__ Wh,
                                                               unsigned short* __restrict__ Wl,
                                                               const float* __restrict__ sc) {
    split_panels_body<H16::F16>(L, b, n, k, np, rs, m, Wh, Wl, sc);
}

// The fp16 split's scale. Every entry of a Cholesky factor of A satisfies
// |L_ij| <= sqrt(A_ii), so one scale serves all of a factorisation's panels:
// hilo_diag_max_kernel takes sqrt of the largest diagonal entry of the
// FP32-stored matrix (rows below N; one work-group per diagonal block),
// hilo_abs_max_kernel the largest |entry| of one panel (block column k, rows
// k+1 .. n-1) for the per-step entry; both leave a non-negative bound in
// sc[2] (as its bits: for non-negative floats the unsigned order is the
// numeric order, so work-groups combine by atomicMax; a negative or NaN
// diagonal counts as 0, and the factorisation reports its pivot). hilo_scale_finish_kernel then writes s = 2^(15 - E)
// with bound = f 2^E, f in [0.5, 1), so that s * bound < 2^15 (the fp16
// range ends at 65504), and sc[1] = 1 / s^2. The exponent shift is clamped to
// [-60, 60] so that 1 / s^2 stays a normal FP32 number.
__global__ void hilo_zero_kernel(float* __restrict__ sc) {
    if (threadIdx.x < 4) sc[threadIdx.x] = 0.0f;
}

// work-group J: the diagonal of block J (after hilo_zero_kernel)
__global__ __launch_bounds__(256) void hilo_diag_max_kernel(const float* __restrict__ L, int b, int n, i64 N,
                                                            float* __restrict__ sc) {
    __shared__ float red[256];
    const int J = blockIdx.x;
    const float* D = L + hs_block_id(J, J, n) * (i64)b * b;
    float mx = 0.0f;
    for (int r = threadIdx.x; r < b && (i64)J * b + r < N; r += 256) mx = fmaxf(mx, D[(i64)r * b + r]);
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(reinterpret_cast<unsigned*>(sc + 2), __float_as_uint(sqrtf(red[0])));
}

__global__ __launch_bounds__(256) void hilo_abs_max_kernel(const float* __restrict__ A, int b, int n, int k,
                                                           float* __restrict__ sc) {
    __shared__ float red[256];
    const i64 bb = (i64)b * b;
    const i64 first = hs_block_id(k + 1, k, n) * bb, total = (i64)(n - k - 1) * bb;
    float mx = 0.0f;
    for (i64 e = (i64)blockIdx.x * 256 + threadIdx.x; e < total; e += (i64)gridDim.x * 256)
        mx = fmaxf(mx, fabsf(A[first + e]));
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(reinterpret_cast<unsigned*>(sc + 2), __float_as_uint(red[0]));
}

__global__ void hilo_scale_finish_kernel(float* __restrict__ sc) {
    if (threadIdx.x != 0) return;
    const float bound = sc[2];
    int e = 0;
    if (bound > 0.0f && bound <= 3.0e38f) frexpf(bound, &e);
    const int sh = min(60, max(-60, 15 - e));
    sc[0] = ldexpf(1.0f, sh);
    sc[1] = ldexpf(1.0f, -2 * sh);
}

}  // namespace

// The dynamic LDS limit of a 256-tile kernel, set on the device current at
// the launch (the attribute belongs to that device).
static int set_lds(const void* fn, size_t bytes) {
    HS_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    return HS_OK;
}

// FP32 updates of several panels, and whole-triangle steps, in 256 x 256
// tiles (HS_F32_TILE256=0 keeps the 128 x 128 kernel)
static const bool g_f32_ps256 = [] {
    const char* e = std::getenv("HS_F32_TILE256");
    return !(e && e[0] == '0');
}();

// planes (hi+lo variants only): the np panels k .. k+np-1 as hi / lo planes
// for the block rows from r0 (split_panels_kernel layout, m = n - r0); null
// for an update that splits the FP32 values it loads. sc: the fp16 split's
// scale (HS_MIXED_F16X3 only).
int hs_launch_trailing_mixed(hs_handle_t h, hipStream_t stream, float* A, int b, int k, int rowStart, int rowCount,
                             int n, int mode, int variant, int np, const unsigned short* planes, int r0,
                             const float* sc) {
    if (rowCount <= 0) return HS_OK;
    if (b % 128 != 0) {
        hs_set_error("mixed-precision trailing update needs b % 128 == 0");
        return HS_ERR_UNSUPPORTED;
    }
    constexpr int TILE = 128;
    const i64 tpb = (i64)(b / TILE) * (b / TILE);
    PairMap pm;
    const i64 grid = hs_pair_map(k, rowStart, rowCount, n, mode, np, tpb, &pm);
    if (grid <= 0) return HS_OK;
    if (grid > 0x7fffffffLL) {
        hs_set_error("trailing update grid too large");
        return HS_ERR_UNSUPPORTED;
    }
    const dim3 g((unsigned)grid), t(THREADS);
    const bool f16 = variant == HS_MIXED_F16X3;
    if (variant == HS_MIXED_BF16X3 || f16) {
        if (planes) {
            const int m = n - r0;
            const unsigned short* Wh = planes;
            const unsigned short* Wl = planes + (i64)np * m * b * b;
            if (b % T2 == 0) {
                PairMap pm2;
                const i64 grid2 = hs_pair_map(k, rowStart, rowCount, n, mode, np, (i64)(b / T2) * (b / T2), &pm2);
                const dim3 g2((unsigned)grid2), t2(T2THREADS);
                if (grid2 > 0 && f16) {
                    if (int rc = set_lds((const void*)trailing_f16x3_ps256, T2LDS)) return rc;
                    trailing_f16x3_ps256<<<g2, t2, T2LDS, stream>>>(A, b, n, pm2, Wh, Wl, r0, m, sc);
                } else if (grid2 > 0) {
                    if (int rc = set_lds((const void*)trailing_bf16x3_ps256, T2LDS)) return rc;
                    trailing_bf16x3_ps256<<<g2, t2, T2LDS, stream>>>(A, b, n, pm2, Wh, Wl, r0, m);
                }
            } else if (f16) {
                if (mode >= 5) trailing_mfma_f16x3<TILE, 1, true><<<g, t, 0, stream>>>(A, b, n, pm, Wh, Wl, r0, m, sc);
                else trailing_mfma_f16x3<TILE, 0, true><<<g, t, 0, stream>>>(A, b, n, pm, Wh, Wl, r0, m, sc);
            } else {
                if (mode >= 5) trailing_mfma_bf16x3<TILE, 1, true><<<g, t, 0, stream>>>(A, b, n, pm, Wh, Wl, r0, m);
                else trailing_mfma_bf16x3<TILE, 0, true><<<g, t, 0, stream>>>(A, b, n, pm, Wh, Wl, r0, m);
            }
        } else if (f16) {
            if (mode >= 5)
                trailing_mfma_f16x3<TILE, 1, false><<<g, t, 0, stream>>>(A, b, n, pm, nullptr, nullptr, 0, 0, sc);
            else
                trailing_mfma_f16x3<TILE, 0, false><<<g, t, 0, stream>>>(A, b, n, pm, nullptr, nullptr, 0, 0, sc);
        } else if (mode >= 5) {
            trailing_mfma_bf16x3<TILE, 1, false><<<g, t, 0, stream>>>(A, b, n, pm, nullptr, nullptr, 0, 0);
        } else {
            trailing_mfma_bf16x3<TILE, 0, false><<<g, t, 0, stream>>>(A, b, n, pm, nullptr, nullptr, 0, 0);
        }
    } else if (b % T2 == 0 && (np > 1 || mode == 3) && g_f32_ps256) {
        PairMap pm2;
        const i64 grid2 = hs_pair_map(k, rowStart, rowCount, n, mode, np, (i64)(b / T2) * (b / T2), &pm2);
        if (grid2 > 0) {
            if (int rc = set_lds((const void*)trailing_f32_ps256, F2LDS)) return rc;
            trailing_f32_ps256<<<dim3((unsigned)grid2), dim3(T2THREADS), F2LDS, stream>>>(A, b, n, pm2);
        }
    } else {
        if (mode >= 5) trailing_mfma_f32<TILE, 1><<<g, t, 0, stream>>>(A, b, n, pm);
        else trailing_mfma_f32<TILE, 0><<<g, t, 0, stream>>>(A, b, n, pm);
    }
    HS_LAUNCH_CHECK();
    return HS_OK;
}

// The fp16 split's scale for a factorisation of the FP32-stored matrix L
// (from its largest diagonal entry), into h->hilo_scale on h->stream.
static int hilo_scale_for_factor(hs_handle_t h, const float* L, int b, int n, i64 N) {
    hilo_zero_kernel<<<1, 64, 0, h->stream>>>(h->hilo_scale);
    HS_LAUNCH_CHECK();
    hilo_diag_max_kernel<<<dim3((unsigned)n), 256, 0, h->stream>>>(L, b, n, N, h->hilo_scale);
    HS_LAUNCH_CHECK();
    hilo_scale_finish_kernel<<<1, 64, 0, h->stream>>>(h->hilo_scale);
    HS_LAUNCH_CHECK();
    return HS_OK;
}

// The fused trailing update of one right-looking step on FP32 storage, in
// any variant (the mixed factor's per-op entry points). The fp16 split takes
// its scale from the step's own panel.
static int trailing_mixed_step(hs_handle_t h, float* A, int b, int k, int blockCountXY, int variant) {
    if (!h) return -1;
    if (!A) return -2;
    if (b <= 0) return -3;
    if (k < 0 || k >= blockCountXY) return -4;
    if (variant == HS_MIXED_F16X3 && b % 128 == 0 && k + 1 < blockCountXY) {
        const i64 total = (i64)(blockCountXY - k - 1) * b * b;
        const unsigned G = (unsigned)std::min<i64>(1024, (total + 4095) / 4096);
        hilo_zero_kernel<<<1, 64, 0, h->stream>>>(h->hilo_scale);
        HS_LAUNCH_CHECK();
        hilo_abs_max_kernel<<<G, 256, 0, h->stream>>>(A, b, blockCountXY, k, h->hilo_scale);
        HS_LAUNCH_CHECK();
        hilo_scale_finish_kernel<<<1, 64, 0, h->stream>>>(h->hilo_scale);
        HS_LAUNCH_CHECK();
    }
    return hs_launch_trailing_mixed(h, h->stream, A, b, k, k + 1, blockCountXY - k - 1, blockCountXY, 3, variant, 1,
                                    nullptr, 0, h->hilo_scale);
}

extern "C" int hs_syrk_gemm_trailing_f32(hs_handle_t h, float* A, int b, int k, int blockCountXY) {
    return trailing_mixed_step(h, A, b, k, blockCountXY, HS_MIXED_FP32);
}

extern "C" int hs_syrk_gemm_trailing_bf16x3(hs_handle_t h, float* A, int b, int k, int blockCountXY) {
    return trailing_mixed_step(h, A, b, k, blockCountXY, HS_MIXED_BF16X3);
}

extern "C" int hs_syrk_gemm_trailing_f16x3(hs_handle_t h, float* A, int b, int k, int blockCountXY) {
    return trailing_mixed_step(h, A, b, k, blockCountXY, HS_MIXED_F16X3);
}

extern "C" int hs_to_f32_packed(hs_handle_t h, const double* K, float* L32, int64_t N, int b, double shift) {
    if (!h) return -1;
    if (!K) return -2;
    if (!L32) return -3;
    if (N <= 0) return -4;
    if (b <= 0) return -5;
    const int n = (int)((N + b - 1) / b);
    const i64 total = (i64)n * (n + 1) / 2 * b * b;
    const int vec = ((uintptr_t)K % 16 == 0 && (uintptr_t)L32 % 16 == 0) ? 1 : 0;
    to_f32_kernel<<<dim3(8192), dim3(256), 0, h->stream>>>(K, L32, total, vec);
    HS_LAUNCH_CHECK();
    if (shift != 0.0) {
        diag_shift_kernel<<<dim3((unsigned)n), dim3(256), 0, h->stream>>>(K, L32, b, n, N, shift);
        HS_LAUNCH_CHECK();
    }
    return HS_OK;
}

// hi / lo planes of the np panels k .. k+np-1 for the m block rows from r0
// (W: hi planes, then lo planes), bf16 or scaled fp16 per variant
static int split_panels(hipStream_t stream, const float* L, int b, int n, int k, int np, int r0, int m,
                        unsigned short* W, int variant, const float* sc) {
    const i64 half = (i64)np * m * b * b;
    const i64 sg = std::min<i64>(8192, (half / 8 + 255) / 256);
    if (variant == HS_MIXED_F16X3)
        split_panels_f16_kernel<<<dim3((unsigned)sg), dim3(256), 0, stream>>>(L, b, n, k, np, r0, m, W, W + half, sc);
    else
        split_panels_bf16_kernel<<<dim3((unsigned)sg), dim3(256), 0, stream>>>(L, b, n, k, np, r0, m, W, W + half);
    HS_LAUNCH_CHECK();
    return HS_OK;
}

// Panels per step of the mixed factor's look-ahead schedule: 6. Every rest-
// update launch reads and writes its FP32 C tiles once, so deeper updates
// spend less of the launch in that epilogue (BF16 hi+lo 256-tile update
// alone: 1165 / 1224 / 1257 TFLOP/s over the three products for 4 / 6 / 8
// panels), against a longer panel chain per step. Factor at N = 65536,
// b = 256 for 4 / 6 / 8 panels: BF16 hi+lo 324 / 318 / 322 ms, FP32 808 / 797
// / 806 ms; with the current kernels BF16 hi+lo 297 / 290 / 291, scaled FP16
// hi+lo 305 / 299 / 298, FP32 744 / 732 / 728 ms. HS_MIXED_PANELS = 1..8
// chooses another depth.
static int mixed_panels(int n, int variant) {
    (void)variant;
    static const int v = [] {
        const char* e = std::getenv("HS_MIXED_PANELS");
        const int w = e ? std::atoi(e) : 0;
        return (w >= 1 && w <= 8) ? w : 6;
    }();
    return std::max(1, std::min(v, n / 2));
}

// the hi+lo factors' two plane sets (the FP32 one needs no workspace)
size_t hs_ws_mixed(i64 N, int b) {
    const int n = (int)((N + b - 1) / b);
    const i64 planeSet = 2 * (i64)mixed_panels(n, HS_MIXED_BF16X3) * n * b * b;
    return sizeof(unsigned short) * 2 * (size_t)planeSet;
}

static int mixed_factor(hs_handle_t h, float* L32, i64 N, int b, int variant, PhaseTimer& pt) {
    const int n = (int)((N + b - 1) / b);
    int rc = hs_reset_info(h);
    if (rc) return rc;
    const int P = mixed_panels(n, variant);
    // hi+lo variants: the P panels of a step as hi / lo planes for the block
    // rows below them (4 bytes per element), split once on the panel stream
    // when the panels are final and read by both the step's column update
    // (S2) and its rest update (S1). Two sets alternate by step: the set of
    // step s is rewritten at the end of step s+1's panel chain, after S2 has
    // waited for S1's update of step s.
    unsigned short* planes = nullptr;
    const i64 planeSet = 2 * (i64)P * n * b * b;
    if (variant == HS_MIXED_BF16X3 || variant == HS_MIXED_F16X3) {
        if ((rc = hs_work(h, sizeof(unsigned short) * 2 * (size_t)planeSet, (void**)&planes))) return rc;
    }
    // the fp16 split's scale, on the ordering stream before the schedule
    // starts (the panel stream waits for everything queued there)
    if (variant == HS_MIXED_F16X3 && (rc = hilo_scale_for_factor(h, L32, b, n, N))) return rc;
    const float* sc = h->hilo_scale;
    auto planesOf = [&](int k) { return planes + ((k / P) & 1) * planeSet; };
    auto prep = [&](hipStream_t s, int k0) -> int {
        if (!planes || !hs_plane_prep_needed(k0, P, n)) return HS_OK;
        return split_panels(s, L32, b, n, k0, P, k0 + P, n - (k0 + P), planesOf(k0), variant, sc);
    };
    auto potrf = [&](int k) { return hs_factor_potrf<float>(h, L32, N, b, hs_block_id(k, k, n), k); };
    auto trsm = [&](int k, int start, int count) {
        return hs_factor_trsm<float>(h, L32, b, hs_block_id(k, k, n), k, start, count);
    };
    auto trail = [&](hipStream_t s, int k, int mode, int rowStart, int rowCount, int np) {
        // the step's rest update and its P-panel column update read the split planes
        const bool shared = planes && hs_trail_reads_planes(mode, np, P);
        return hs_launch_trailing_mixed(h, s, L32, b, k, rowStart, rowCount, n, mode, variant, np,
                                        shared ? planesOf(k) : nullptr, k + P, sc);
    };
    if ((rc = hs_lookahead_factor(h, n, pt, potrf, trsm, trail, P, prep))) {
        hipStreamSynchronize(h->stream2);
        return rc;
    }
    int64_t info = 0;
    if ((rc = hs_get_info(h, &info))) return rc;
    return (int)info;
}

// Work of the rest-update launches of the mixed factor's schedule for N, b
// and variant: their count and algorithmic flops (one product per element).
extern "C" int hs_cholesky_mixed_schedule(int64_t N, int b, int variant, int64_t* rest_launches, double* rest_flops) {
    if (N <= 0) return -1;
    if (b % 128 != 0 || b > 256) return -2;
    if (variant != HS_MIXED_FP32 && variant != HS_MIXED_BF16X3 && variant != HS_MIXED_F16X3) return -3;
    if (!rest_launches) return -4;
    if (!rest_flops) return -5;
    const int n = (int)((N + b - 1) / b);
    hs_rest_update_work(n, b, rest_launches, rest_flops, mixed_panels(n, variant));
    return HS_OK;
}

extern "C" int hs_cholesky_mixed(hs_handle_t h, const double* K, float* L32, int64_t N, int b, int variant,
                                 double shift) {
    if (!h) return -1;
    if (!K) return -2;
    if (!L32) return -3;
    if (N <= 0) return -4;
    if (b % 128 != 0 || b > 256) {
        hs_set_error("hs_cholesky_mixed: matrix_bsz must be 128 or 256");
        return -5;
    }
    if (variant != HS_MIXED_FP32 && variant != HS_MIXED_BF16X3 && variant != HS_MIXED_F16X3) return -6;
    int rc = hs_to_f32_packed(h, K, L32, N, b, shift);
    if (rc) return rc;
    PhaseTimer pt;
    return mixed_factor(h, L32, N, b, variant, pt);
}

extern "C" int hs_cholesky_mixed_timed(hs_handle_t h, const double* K, float* L32, int64_t N, int b, int variant,
                                       double shift, double* phase_ms, int64_t* launches) {
    if (!phase_ms) return -8;
    if (!launches) return -9;
    if (!h) return -1;
    if (!K) return -2;
    if (!L32) return -3;
    if (N <= 0) return -4;
    if (b % 128 != 0 || b > 256) {
        hs_set_error("hs_cholesky_mixed: matrix_bsz must be 128 or 256");
        return -5;
    }
    if (variant != HS_MIXED_FP32 && variant != HS_MIXED_BF16X3 && variant != HS_MIXED_F16X3) return -6;
    int rc = hs_to_f32_packed(h, K, L32, N, b, shift);
    if (rc) return rc;
    PhaseTimer pt;
    pt.on = true;
    rc = mixed_factor(h, L32, N, b, variant, pt);
    if (rc < 0) return rc;
    int rc2 = pt.collect(phase_ms, launches);
    return rc2 ? rc2 : rc;
}
