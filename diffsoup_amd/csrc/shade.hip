// Fused per-pixel shading for the training step: multires feature gather ->
// pixel MLP (Linear-ReLU-Linear-ReLU-Linear-Sigmoid) -> residual blend ->
// MSE against the target, and the whole adjoint (MLP input and weight
// gradients, feature-lattice scatter) in the same pass.
//
// Work unit: the covered pixels, grouped by triangle, are split into equal
// contiguous ranges, one per wave.  For each triangle segment in its range a
// wave stages the triangle's feature row (all levels, all channels) in LDS,
// shades the segment's pixels (any view) in tiles of 16, accumulates the input
// gradient into an LDS gradient row (lattice levels 0-2 as fp32 MFMA products
// accumulated in registers over the segment, finer levels by claimed LDS
// read-modify-writes), and writes that row once - with plain coalesced stores when the
// triangle lies wholly in the range, with global atomics only for the few
// triangles split between two ranges.  Every layer is an MFMA product with
// the layer's output units on the M axis and the 16 pixels on the N axis.  In
// the accumulator layout of v_mfma_f32_16x16x{4f32,16bf16} lane l holds rows
// 4(l/16) + r, r = 0..3, of pixel l % 16, which is exactly the B-operand
// layout the next layer needs (with the K order of the f32 variant permuted),
// so activations stay in registers from the feature gather to the MLP input
// gradient.  Weights are pre-arranged per lane ("blobs") once per step and
// staged in LDS.  Weight gradients contract over pixels, so the per-pixel
// activations and deltas are transposed through a per-wave LDS tile and fed
// to MFMAs whose accumulators persist across all tiles a wave shades; one
// partial per workgroup is summed by the step's epilogue kernel.
//
// Reference behaviour: ColorMLP (python/diffsoup/multires.py:109-207),
// multires_triangle_color (src/cuda/multires.cu:224-312, here per level
// concatenated), the training-step composition of
// examples/02_synthetic.py:354-388 with an MSE loss, and the feature scatter
// of src/cuda/multires.cu:70-128.
#include "ds_kernels.h"
#include "ds_math.h"
#include "ds_shade.h"

#include <algorithm>
#include <type_traits>

namespace ds {
namespace {

constexpr int kWave = 64;
constexpr int kWavesPerBlock = kShadeWavesPerBlock;
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef short i16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef short i16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// Four fp32 values to bf16, rounded to nearest even, by the hardware's
// packed conversion (v_cvt_pk_bf16_f32, two values per instruction).
__device__ __forceinline__ i16x4 to_bf16x4(f32x4 v) {
    return __builtin_bit_cast(i16x4, __builtin_convertvector(v, bf16x4));
}

__device__ __forceinline__ float sigmoidf(float z) { return 1.0f / (1.0f + expf(-z)); }

// Sum over the 16 lanes of this lane's DPP row (lanes 16 r .. 16 r + 15),
// every lane receiving the total: four row rotations (row_ror 8, 4, 2, 1)
// on the VALU instead of LDS permutes.  All lanes must be active.
__device__ __forceinline__ float row16_sum(float x) {
    x += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x128, 0xf, 0xf, false));
    x += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x124, 0xf, 0xf, false));
    x += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x122, 0xf, 0xf, false));
    x += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x121, 0xf, 0xf, false));
    return x;
}


// Compile-time shape of one shading configuration: NL levels of C channels
// (concatenated), optionally followed by the 9 SH2 values of the pixel's
// view direction (SH), into an MLP of hidden width DH.  In LDS a sample's
// channels are padded to CP, a multiple of 4, so every 4-channel block is one
// 16-byte access; a channel count that is not a multiple of 4 needs a single
// level (the reference's accumulated features: 7 channels + SH2 = 16 inputs).
template <int NL_, int C_, int DH_, bool BF16_, bool SH_ = false>
struct Cfg {
    static constexpr int NL = NL_, C = C_, DH = DH_;
    static constexpr bool BF16 = BF16_, SH = SH_;
    static constexpr int CP = (C + 3) & ~3;
    static constexpr int DF = NL * C;            // feature inputs
    static constexpr int DIN = DF + (SH ? 9 : 0);
    static constexpr int NI = (DIN + 15) / 16;  // input tiles
    static constexpr int NH = DH / 16;          // hidden tiles
    static constexpr int DINP = 16 * NI;
    // blob layer table: (out tiles, in tiles) for L1, L2, L3, L3T, L2T, L1T
    static constexpr int L1 = 0, L2 = NH * NI, L3 = L2 + NH * NH, L3T = L3 + NH, L2T = L3T + NH,
                         L1T = L2T + NH * NH, NTILES = L1T + NI * NH;
    // operand units of one (m, t) tile pair, in floats: 4 f32 steps of 64
    // lanes, or one bf16 step of 64 lanes x 4 bf16 (128 floats' worth)
    static constexpr int TILE_ELEMS = (BF16 ? 2 : 4) * kWave;
    static constexpr int BLOB_ELEMS = NTILES * TILE_ELEMS;
    static constexpr int NPARAM = DH * DIN + DH * DH + 3 * DH + DH + DH + 3 + 1;
    // workgroups per CU the register allocation must allow: two (two waves
    // per SIMD, 256 registers each) up to hidden width 32; the 80->64->64->3
    // configuration needs one wave's full 512 registers
    static constexpr int MIN_BLOCKS = DH <= 32 ? 2 : 1;
    // claim rounds of the feature-gradient scatter before the remaining
    // conflicts fall back to LDS atomics: a second round settles most of the
    // conflicts at the coarser fine levels (eight 800x800 views, levels 0-3:
    // shading 169 -> 160 us) but costs the 80->64->64->3 configuration
    // registers it does not have (2.1 -> 3.0 ms)
    static constexpr int CLAIM_ROUNDS = 2;
    // one wave per SIMD (512 registers): lattice levels 0-2 scatter by fp32
    // MFMA, with their accumulators and the bias gradients' block in
    // registers (80->64->64->3 on 100K triangles: shading 2.13 -> 1.77 ms).
    // At two waves per SIMD the per-lane coarse sums, claims and per-lane
    // bias sums are faster (eight 800x800 views, levels 0-3, 32-wide MLP:
    // 159 us against 177-192 us with the MFMA forms): fp32 MFMA issue is
    // what the smaller networks are short of.
    static constexpr bool MF = MIN_BLOCKS == 1;
    // a triangle's segment cost, in pixels, in the waves' range split (row
    // staging, coarse reduction, row store; where tiles stop at a triangle's
    // end, also half a tile of idle lanes).  Measured: 6 / 8 / 10 / 12 / 16 /
    // 24 give 145 / 141 / 138 / 138 / 141 / 147 us on eight 800x800 views
    // (tiles crossing triangles), and 0 / 2 / 4 / 8 / 16 give 1.61 / 1.49 /
    // 1.48 / 1.51 / 1.55 ms for the 80->64->64->3 kernel on 100K triangles
    static constexpr long long SEG_COST = MF ? 4 : 12;
    // per-wave LDS staging for the weight-gradient transposes: one (delta,
    // activation) pair at a time - (D1, X), (D2, H1), (D3, H2).  The f32
    // products stage floats as [pixel][unit] rows (width + 4 apart); the bf16
    // products stage bf16 values (their MFMA operands are rounded to bf16
    // anyway) as [unit][pixel] rows of SPX, so that a lane's four K values
    // (pixels 4g .. 4g+3) are one 8-byte read.
    using ST = typename std::conditional<BF16, unsigned short, float>::type;
    static constexpr int SPX = 20;
    static constexpr int region(int width) { return BF16 ? width * SPX : 16 * (width + 4); }
    static constexpr int STAGE_ = region(DH) + region(DINP) > 2 * region(DH) ? region(DH) + region(DINP)
                                                                              : 2 * region(DH);
    static constexpr int STAGE = STAGE_ > region(16) + region(DH) ? STAGE_ : region(16) + region(DH);
    static_assert(C % 4 == 0 || NL == 1, "channels per level must be a multiple of 4 with several levels");
    static_assert(NTILES == blob_tiles(NI, NH) && TILE_ELEMS == blob_tile_floats(BF16),
                  "the step prologue's blob layout (BlobJob) must be this one");
    static_assert(DH % 16 == 0, "hidden width must be a multiple of 16");
};

constexpr int up16(int b) { return (b + 15) & ~15; }

// Floats per row slot of the visible-opacity rows: the triangle's Sa opacity
// values, then their Sa gradients, padded to whole float4s.
__host__ __device__ constexpr int va_floats(int Sa) { return (2 * Sa + 3) & ~3; }

// Lattice levels up to this one (3, 6 and 15 samples: one 16-row MFMA block
// each) scatter their feature gradient by MFMA; finer levels by claims.
constexpr int kMfmaTopLevel = 2;
// What that scatter stages, one level at a time, in the per-wave stage
// region: the level's input tile of dX transposed ([16 pixels][16] floats,
// rows 20 apart) and its 16 pixels' corner records (three indices, three
// weights: 32 B each).
template <class K>
struct MfPlan {
    static constexpr int NMF = K::NL < kMfmaTopLevel + 1 ? K::NL : kMfmaTopLevel + 1;
    static constexpr int BYTES = 16 * 20 * 4 + 16 * 32;
};

// Feature-row length (floats) of levels 0 .. nl-1 with c channels.
constexpr int row_floats_from_zero(int nl, int c) {
    int s = 0;
    for (int l = 0; l < nl; ++l) s += l == 0 ? 3 : ((1 << (l - 1)) + 1) * ((1 << l) + 1);
    return s * c;
}

// LDS plan of one configuration and feature-row precision: the weight
// operands, then per wave the feature row (kept in bf16 when the features are
// given in bf16 - they are widened exactly on use), the fp32 gradient row and
// the transpose stage.  The weight operands go to LDS when the whole plan
// fits in 160 KB for rows starting at level 0; otherwise they are read from
// global memory, where they stay L2-resident.
template <class K, bool FB>
struct Plan {
    using FT = typename std::conditional<FB, unsigned short, float>::type;
    __host__ __device__ static constexpr int frow_bytes(int rowp) { return up16(rowp * static_cast<int>(sizeof(FT))); }
    __host__ __device__ static constexpr int stage_bytes() {
        return !K::MF || up16(K::STAGE * static_cast<int>(sizeof(typename K::ST))) > MfPlan<K>::BYTES
                   ? up16(K::STAGE * static_cast<int>(sizeof(typename K::ST)))
                   : MfPlan<K>::BYTES;
    }
    // NS slots (two where tiles run on into the next triangle: the kernels
    // without MF) of feature row and gradient row, the transpose stage, and
    // (CLAIM) NS claim tables of the gradient rows' 4-channel blocks (one int
    // each: rowp bytes)
    static constexpr int NS = K::MF ? 1 : 2;
    __host__ __device__ static constexpr int base_bytes(int rowp, int ns = NS) {
        return ns * (frow_bytes(rowp) + 4 * rowp) + stage_bytes();
    }
    static constexpr int ROW0 = (row_floats_from_zero(K::NL, K::CP) + 3) & ~3;
    static constexpr int LDS_MAX = 160 * 1024, BLOB = 4 * K::BLOB_ELEMS;
    // preference, for rows starting at level 0: operands and claim table in
    // LDS; claim table only; operands only; neither
    static constexpr bool FIT_BOTH = BLOB + kWavesPerBlock * (base_bytes(ROW0, NS) + NS * ROW0) <= LDS_MAX;
    static constexpr bool FIT_CLAIM = kWavesPerBlock * (base_bytes(ROW0, NS) + NS * ROW0) <= LDS_MAX;
    static constexpr bool FIT_BLOB = BLOB + kWavesPerBlock * base_bytes(ROW0, NS) <= LDS_MAX;
    static constexpr bool CLAIM = FIT_BOTH || FIT_CLAIM;
    static constexpr bool LB = FIT_BOTH || (!FIT_CLAIM && FIT_BLOB);
    static constexpr int BLOB_BYTES = LB ? BLOB : 0;
    __host__ __device__ static constexpr int wave_bytes(int rowp, int ns = NS) {
        return base_bytes(rowp, ns) + (CLAIM ? ns * rowp : 0);
    }
};

__device__ __forceinline__ float4 ld4(const float* p) { return *reinterpret_cast<const float4*>(p); }
__device__ __forceinline__ float4 ld4(const unsigned short* p) {
    const uint2 u = *reinterpret_cast<const uint2*>(p);
    return make_float4(__uint_as_float(u.x << 16), __uint_as_float(u.x & 0xffff0000u),
                       __uint_as_float(u.y << 16), __uint_as_float(u.y & 0xffff0000u));
}
// Stage lane (g, j)'s four values (units u0 .. u0+3 of pixel j) into a
// transpose region of `width` units.
template <class K>
__device__ __forceinline__ void stage_put(typename K::ST* reg, int width, int u0, int j, f32x4 v) {
    if constexpr (K::BF16) {
        const i16x4 b = to_bf16x4(v);
        for (int r = 0; r < 4; ++r) reg[(u0 + r) * K::SPX + j] = static_cast<unsigned short>(b[r]);
    } else {
        *reinterpret_cast<float4*>(reg + j * (width + 4) + u0) = make_float4(v[0], v[1], v[2], v[3]);
    }
}

// LDS slot of channels c .. c+3 (c a multiple of 4) of sample s in a
// gradient row.  The row keeps the table's sample-major layout, but the
// 4-channel blocks of a sample are permuted by an XOR with bits of s, so
// that the pixel lanes of one atomic instruction (which add the same channel
// block to different lattice samples) spread over all 32 LDS banks instead
// of 32 / C * 4 of them.
template <int C>
__device__ __forceinline__ int grad_slot(uint32_t s, int c) {
    constexpr int NB = C / 4;
    if constexpr ((NB & (NB - 1)) == 0 && C <= 32) {
        const int sw = static_cast<int>((s * C) >> 5) & (NB - 1);
        return static_cast<int>(s) * C + 4 * ((c >> 2) ^ sw);
    } else {
        return static_cast<int>(s) * C + c;
    }
}

// One MFMA step: acc += A * B on a 16x16 tile.  For f32 an operand unit is a
// single float (K = 4 per step over 4 lane groups), for bf16 four values
// (K = 16 per step).
template <bool BF16>
struct Mma;

template <>
struct Mma<false> {
    // in: the input tile t of the chained activations (D layout, 4 floats)
    // blob: per-lane values for (m, t): 4 consecutive steps r, 64 lanes each
    __device__ static __forceinline__ f32x4 tile(const float* blob, int lane, f32x4 in, f32x4 acc) {
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(blob[0 * kWave + lane], in[0], acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(blob[1 * kWave + lane], in[1], acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(blob[2 * kWave + lane], in[2], acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(blob[3 * kWave + lane], in[3], acc, 0, 0, 0);
        return acc;
    }
    // weight-gradient step over 16 pixels: A[i][k] = a_lds(k = pixel, i), B[k][j] = b_lds(pixel, j)
    __device__ static __forceinline__ f32x4 outer(const float* A, int sa, const float* Bm, int sb, int lane,
                                                  f32x4 acc) {
        const int i = lane % 16, g = lane / 16;
        for (int s = 0; s < 4; ++s) {
            const int px = 4 * s + g;
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(A[px * sa + i], Bm[px * sb + i], acc, 0, 0, 0);
        }
        return acc;
    }
};

template <>
struct Mma<true> {
    __device__ static __forceinline__ i16x4 pack(f32x4 v) { return to_bf16x4(v); }
    // in: the input tile already packed to bf16
    __device__ static __forceinline__ f32x4 tile_packed(const float* blob, int lane, i16x4 in, f32x4 acc) {
        const i16x4 a = reinterpret_cast<const i16x4*>(blob)[lane];
        return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, in, acc, 0, 0, 0);
    }
    // Two adjacent input tiles t, t+1 in one 16x16x32 product: lane (i, g)
    // supplies K slots 8g .. 8g+7 as tile t's four values (units 4g .. 4g+3)
    // followed by tile t+1's; A and B use the same slot order, so the sum
    // over K is the two 16x16x16 products' sum.
    __device__ static __forceinline__ f32x4 tile_pair(const float* blob0, const float* blob1, int lane, i16x4 in0,
                                                      i16x4 in1, f32x4 acc) {
        const i16x4 a0 = reinterpret_cast<const i16x4*>(blob0)[lane];
        const i16x4 a1 = reinterpret_cast<const i16x4*>(blob1)[lane];
        const i16x8 a = __builtin_shufflevector(a0, a1, 0, 1, 2, 3, 4, 5, 6, 7);
        const i16x8 b = __builtin_shufflevector(in0, in1, 0, 1, 2, 3, 4, 5, 6, 7);
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b),
                                                       acc, 0, 0, 0);
    }
    // staged bf16 values, [unit][pixel] rows of spx: A(i, k) = At[i][k], B(k, j) = Bt[j][k]
    __device__ static __forceinline__ f32x4 outer(const unsigned short* At, const unsigned short* Bt, int spx,
                                                  int lane, f32x4 acc) {
        const int i = lane % 16, g = lane / 16;
        const i16x4 a = *reinterpret_cast<const i16x4*>(At + i * spx + 4 * g);
        const i16x4 b = *reinterpret_cast<const i16x4*>(Bt + i * spx + 4 * g);
        return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
};

// acc[i][n] += sum over the 16 staged pixels of unit ua + i of transpose
// region A (width wa), for column n == col only: the B operand is that
// column's indicator.  Bias gradients contract a delta with ones.
template <class K>
__device__ __forceinline__ f32x4 stage_sum(const typename K::ST* A, int wa, int ua, int col, int lane, f32x4 acc) {
    const int i = lane % 16, g = lane / 16;
    if constexpr (K::BF16) {
        const i16x4 a = *reinterpret_cast<const i16x4*>(A + (ua + i) * K::SPX + 4 * g);
        const short one = i == col ? static_cast<short>(0x3f80) : static_cast<short>(0);   // bf16 1.0
        const i16x4 b = {one, one, one, one};
        return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    } else {
        const float one = i == col ? 1.f : 0.f;
        for (int s = 0; s < 4; ++s)
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(A[(4 * s + g) * (wa + 4) + ua + i], one, acc, 0, 0, 0);
        return acc;
    }
}

// Offset (in floats) of a layer tile pair inside the blob table.
template <class K>
__device__ __forceinline__ const float* blob_at(const float* blobs, int layer_base, int m, int nin, int t) {
    return blobs + (layer_base + m * nin + t) * K::TILE_ELEMS;
}

// acc += (delta units ua .. ua+15) x (activation units ub .. ub+15)^T over
// the 16 staged pixels; A, B are transpose regions of widths wa, wb.
template <class K>
__device__ __forceinline__ f32x4 stage_outer(const typename K::ST* A, int wa, int ua, const typename K::ST* B,
                                             int wb, int ub, int lane, f32x4 acc) {
    if constexpr (K::BF16)
        return Mma<true>::outer(A + ua * K::SPX, B + ub * K::SPX, K::SPX, lane, acc);
    else
        return Mma<false>::outer(A + ua, wa + 4, B + ub, wb + 4, lane, acc);
}

// y[m] = bias + sum_t W(m, t) * in[t] for NOUT output tiles and NIN input
// tiles (bf16: each input tile is converted once, not once per output tile).
template <class K, int NOUT, int NIN>
__device__ __forceinline__ void dense(const float* blobs, int base, int lane, const f32x4 (&in)[NIN],
                                      f32x4 (&out)[NOUT]) {
    if constexpr (K::BF16) {
        i16x4 pin[NIN];
        for (int t = 0; t < NIN; ++t) pin[t] = Mma<true>::pack(in[t]);
        for (int m = 0; m < NOUT; ++m) {
            int t = 0;
            for (; t + 1 < NIN; t += 2)
                out[m] = Mma<true>::tile_pair(blob_at<K>(blobs, base, m, NIN, t), blob_at<K>(blobs, base, m, NIN, t + 1),
                                              lane, pin[t], pin[t + 1], out[m]);
            if (t < NIN) out[m] = Mma<true>::tile_packed(blob_at<K>(blobs, base, m, NIN, t), lane, pin[t], out[m]);
        }
    } else {
        for (int m = 0; m < NOUT; ++m)
            for (int t = 0; t < NIN; ++t)
                out[m] = Mma<false>::tile(blob_at<K>(blobs, base, m, NIN, t), lane, in[t], out[m]);
    }
}

// The 4 channels c .. c+3 of one sample straight from global memory (DIRECT
// rows); channels past C read as 0.
template <int C>
__device__ __forceinline__ float4 ld4_global(const float* p, int c) {
    if constexpr (C % 4 == 0) {
        return *reinterpret_cast<const float4*>(p);
    } else {
        return make_float4(c < C ? p[0] : 0.f, c + 1 < C ? p[1] : 0.f, c + 2 < C ? p[2] : 0.f,
                           c + 3 < C ? p[3] : 0.f);
    }
}

// dst[i] = src[i] for i < n over the lanes of one wave, 8 loads per lane in
// flight before the stores.
template <class T>
__device__ __forceinline__ void copy_batched(T* dst, const T* __restrict__ src, int n, int lane) {
    constexpr int U = 8;
    int i = lane;
    for (; i + (U - 1) * kWave < n; i += U * kWave) {
        T v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = src[i + u * kWave];
#pragma unroll
        for (int u = 0; u < U; ++u) dst[i + u * kWave] = v[u];
    }
    for (; i < n; i += kWave) dst[i] = src[i];
}

// e[k] for a lane-varying k in [0, 9) without dynamic register indexing.
// Each value passes through an empty asm first: a chain of selects between
// loads of one array would otherwise be folded back into one load at a
// lane-varying address, and the array kept in scratch memory.
__device__ __forceinline__ float pick9(const float (&e)[9], int k) {
    float v = e[0];
    asm("" : "+v"(v));
#pragma unroll
    for (int i = 1; i < 9; ++i) {
        float ei = e[i];
        asm("" : "+v"(ei));
        v = k == i ? ei : v;
    }
    return v;
}

// Adds the MFMA scatter blocks of the levels 0 .. n-1 of the row (n = the
// number of levels at or below kMfmaTopLevel) into the LDS gradient row.
// Lane (g, n) holds G[4 g + r][n] of block q: sample 4 g + r of level q,
// feature column n of the level's input tile; each (sample, channel) of these
// levels is held by exactly one lane and added by nothing else, so a plain
// read-modify-write.
template <class K, int NMF>
__device__ __forceinline__ void add_mf_blocks(const f32x4 (&acc)[NMF], int n, int lo, const uint32_t* loff,
                                              float* grow, int g, int j) {
    constexpr int C = K::C, CP = K::CP;
#pragma unroll
    for (int q = 0; q < NMF; ++q) {
        if (q >= n) break;
        const int tq = q * C / 16, c0 = q * C - 16 * tq;
        const int S = static_cast<int>(feats_at_level(lo + q));
        const int ch = j - c0;
        if (ch >= 0 && ch < C)
            for (int r = 0; r < 4; ++r) {
                const int sm = 4 * g + r;
                if (sm < S) grow[grad_slot<CP>(loff[q] + sm, ch & ~3) + (ch & 3)] += acc[q][r];
            }
    }
}

// SPLIT selects the two-pass forms (a.fwd_only, a.ext_grad); the fused
// single-pass instance carries neither test.  DIRECT rows are not staged: a
// triangle's row that does not fit the LDS plan (single-level rows at high
// levels, e.g. 561 samples at level 5) is gathered from global memory, where
// a wave's consecutive pixels of one triangle hit the same cache lines, and
// its gradient is added with global atomics into the zeroed table.
template <class K, bool FB, bool SPLIT, bool DIRECT>
__global__ void __launch_bounds__(kWave * kWavesPerBlock, K::MIN_BLOCKS)
k_shade(ShadeArgs a, int rowp, int ns) {
    constexpr int NI = K::NI, NH = K::NH, C = K::C, CP = K::CP, DF = K::DF, DIN = K::DIN, DH = K::DH;
    constexpr bool MF = K::MF;
    constexpr int NMF = MF ? MfPlan<K>::NMF : 1;   // levels that may scatter by MFMA
    static_assert(!FB || C == CP, "bf16 rows need whole 4-channel blocks");
    static_assert(!DIRECT || K::NL == 1, "direct rows are compiled for single-level features");
    using P = Plan<K, FB>;
    using FT = typename P::FT;
    using ST = typename K::ST;
    extern __shared__ float smem[];
    const float* blobs = P::LB ? smem : a.blobs;
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int g = lane / 16, j = lane % 16;
    // VA: the visible fragments' opacity gradient (the fused single pass;
    // see ShadeArgs), in NS slots of va_floats(Sa) after the claim tables.
    // Compiled for the two-waves-per-SIMD kernels only: the one-wave kernel
    // has no LDS margin for the rows beside its weight operands, and with the
    // operands in L2 instead it spills.
    constexpr bool VA_ON = !SPLIT && !DIRECT && !K::MF;
    const bool VA = VA_ON && a.alpha != nullptr;   // wave-uniform
    const int Sa = a.Sa;
    const int vaf = VA ? va_floats(Sa) : 0;
    unsigned char* wbase =
        reinterpret_cast<unsigned char*>(smem) + P::BLOB_BYTES + wave * (P::wave_bytes(rowp, ns) + 4 * ns * vaf);
    // per wave: NS slots of feature row and gradient row, the transpose
    // stage, NS claim tables
    FT* frow0 = reinterpret_cast<FT*>(wbase);
    FT* frow1 = reinterpret_cast<FT*>(wbase + (ns - 1) * P::frow_bytes(rowp));
    float* grow0 = reinterpret_cast<float*>(wbase + ns * P::frow_bytes(rowp));
    float* grow1 = grow0 + (ns - 1) * rowp;
    ST* stage = reinterpret_cast<ST*>(wbase + ns * (P::frow_bytes(rowp) + 4 * rowp));
    int* claim0 = reinterpret_cast<int*>(wbase + P::base_bytes(rowp, ns));
    int* claim1 = claim0 + (ns - 1) * (rowp / 4);
    float* va0 = reinterpret_cast<float*>(wbase + P::wave_bytes(rowp, ns));
    float* va1 = va0 + (ns - 1) * vaf;
    const int row = a.S_tot * C;

    // stage the weight blobs
    if constexpr (P::LB) {
        const float4* src = reinterpret_cast<const float4*>(a.blobs);
        float4* dst = reinterpret_cast<float4*>(smem);
        for (int i = threadIdx.x; i < K::BLOB_ELEMS / 4; i += blockDim.x) dst[i] = src[i];
    }
    __syncthreads();

    // weight-gradient accumulators; the bias gradients either per lane (gb*,
    // reduced over the pixel lanes at the end) or (MF) in one MFMA block
    // gbias: column m < NH holds b1's units 16 m + row, column NH + m b2's,
    // column 2 NH (rows 0-2) b3's
    f32x4 gW1[NH][NI], gW2[NH][NH], gW3[NH], gb1[NH], gb2[NH], gb3, gbias;
    for (int m = 0; m < NH; ++m) {
        for (int n = 0; n < NI; ++n) gW1[m][n] = f32x4{0.f, 0.f, 0.f, 0.f};
        for (int n = 0; n < NH; ++n) gW2[m][n] = f32x4{0.f, 0.f, 0.f, 0.f};
        gW3[m] = gb1[m] = gb2[m] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
    gb3 = gbias = f32x4{0.f, 0.f, 0.f, 0.f};
    float loss = 0.f;

    // level offsets inside a triangle's feature row
    uint32_t loff[K::NL];
    {
        uint32_t o = 0;
        for (int k = 0; k < K::NL; ++k) { loff[k] = o; o += feats_at_level(a.lo + k); }
    }

    // Each wave owns a contiguous range [bounds[wid], bounds[wid + 1]) of the
    // triangle-grouped pixel list (the ShadeRanges jobs) and walks it in
    // tiles of 16 pixels.  A triangle whose pixels all fall in one wave's
    // range gets its gradient row stored; a triangle split between waves is
    // added with atomics into its row, which the list kernel cleared.  A
    // triangle's end and the next triangle come from records (wave-uniform
    // loads), so the walk never reads the offset table.  XT (two waves per
    // SIMD, rows staged in LDS): a tile may run on into the next triangle,
    // whose row is then staged in the wave's second slot, so that the MLP
    // lanes stay full (on eight 800x800 views 73.5% of the lanes held a
    // pixel when tiles stopped at each triangle's end); the wave's current
    // triangle is A, in slot sA, the next is N.
    const bool XT = !K::MF && !DIRECT && ns == 2;   // wave-uniform
    const int wid = blockIdx.x * kWavesPerBlock + wave;
    const int r0 = a.bounds[wid], r1 = a.bounds[wid + 1];
    const float4* recs4 = reinterpret_cast<const float4*>(a.list);
    // PF: prefetch registers for the next triangle's bf16 feature row (rows up
    // to the plan's level-0 row; longer rows are copied when staged)
    constexpr bool PF = FB && K::MF;
    constexpr int PF_N = PF ? (P::ROW0 / 4 + kWave - 1) / kWave : 1;
    uint2 pf[PF_N];
    bool have_pf = false;

    // Stage triangle t's feature row into slot sl and clear that slot's
    // gradient row.  The row's loads go out in batches of 8 per lane before
    // the first LDS store waits on them: a row is read once per step, so each
    // load is an HBM miss, and a load-store-load chain would pay that latency
    // for every 64 x 8 or 64 x 16 bytes of the row.
    auto begin_row = [&](int sl, int t) {
        FT* frow = sl ? frow1 : frow0;
        float4* g4 = reinterpret_cast<float4*>(sl ? grow1 : grow0);
        if constexpr (VA_ON) {
            if (VA) {
                // the triangle's opacity values and a cleared gradient row
                float* av = sl ? va1 : va0;
                const float* src = a.alpha + (long long)t * Sa;
                for (int i = lane; i < Sa; i += kWave) {
                    av[i] = src[i];
                    av[Sa + i] = 0.f;
                }
            }
        }
        for (int i = lane; i < rowp / 4; i += kWave) g4[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        if constexpr (DIRECT) {
            // nothing staged
        } else if constexpr (C != CP) {
            // channels padded to CP in LDS (the padding stays zero)
            const float* src = a.feat + (long long)t * row;
            for (int i = lane; i < a.S_tot * CP; i += kWave) {
                const int sm = i / CP, c = i - sm * CP;
                frow[i] = c < C ? src[sm * C + c] : 0.f;
            }
        } else if constexpr (FB) {
            // bf16 features stay bf16 in the row: 4 values per 8-byte copy
            // (or, PF, the row the previous triangle's end already loaded)
            if (PF && have_pf) {
#pragma unroll
                for (int u = 0; u < PF_N; ++u)
                    if (lane + u * kWave < row / 4) reinterpret_cast<uint2*>(frow)[lane + u * kWave] = pf[u];
                have_pf = false;
            } else {
                copy_batched(reinterpret_cast<uint2*>(frow),
                             reinterpret_cast<const uint2*>(a.feat_bf16 + (long long)t * row), row / 4, lane);
            }
        } else if (a.feat_bf16) {
            float4* f4 = reinterpret_cast<float4*>(frow);
            // bf16 features: 4 values per 8-byte load, widened exactly to fp32
            const uint2* src = reinterpret_cast<const uint2*>(a.feat_bf16 + (long long)t * row);
            for (int i = lane; i < row / 4; i += kWave) {
                const uint2 u = src[i];
                f4[i] = make_float4(__uint_as_float(u.x << 16), __uint_as_float(u.x & 0xffff0000u),
                                    __uint_as_float(u.y << 16), __uint_as_float(u.y & 0xffff0000u));
            }
        } else {
            copy_batched(reinterpret_cast<float4*>(frow), reinterpret_cast<const float4*>(a.feat + (long long)t * row),
                         row / 4, lane);
        }
        wave_lds_sync();
    };

    // MF: the MFMA-scattered levels' gradient blocks (samples x feature
    // columns of the level's input tile), accumulated over a triangle's
    // tiles; otherwise the two coarsest levels' per-lane sums (racc) of the
    // current triangle's pixels
    f32x4 sacc[NMF];
    for (int q = 0; q < NMF; ++q) sacc[q] = f32x4{0.f, 0.f, 0.f, 0.f};
    float racc[MF ? 1 : 6][4];
    for (int vtx = 0; vtx < (MF ? 1 : 6); ++vtx)
        for (int r = 0; r < 4; ++r) racc[vtx][r] = 0.f;

    int tA = 0, teA = 0, sA = 0, tN = 0, teN = 0;
    bool ownA = false, haveN = false;   // haveN: XT, N staged in slot 1 - sA
    if (r0 < r1) {
        tA = a.list[r0].tri;
        teA = a.list[r0].seg_end;
        ownA = (r0 == 0 || a.list[r0 - 1].tri != tA) && teA <= r1;
        if (teA < r1) {
            tN = a.list[teA].tri;
            teN = a.list[teA].seg_end;
        }
        begin_row(sA, tA);
    }
    // lanes of group 0 also take the pixel's target colour from the record
    float4 rec = r0 + j < r1 ? recs4[2 * (r0 + j)] : make_float4(0.f, 0.f, 0.f, 0.f);
    float4 rtg = g == 0 && r0 + j < r1 ? recs4[2 * (r0 + j) + 1] : make_float4(0.f, 0.f, 0.f, 0.f);
    for (int pos = r0; pos < r1;) {
        // this tile: up to 16 pixels, within A, or (XT) running on into N
        int end = min(pos + 16, r1);
        if (end > teA) {
            if (XT) {
                if (!haveN) {
                    begin_row(1 - sA, tN);
                    haveN = true;
                }
                end = min(end, teN);
            } else {
                end = teA;
            }
        }
        {
            const int idx = pos + j;
            const bool valid = idx < end;
            const bool inA = idx < teA;
            const int sl = inA ? sA : 1 - sA;
            const int tl = inA ? tA : tN;
            const FT* frow_l = sl ? frow1 : frow0;
            float* grow_l = sl ? grow1 : grow0;
            int* claim_l = sl ? claim1 : claim0;
            (void)tl;
            (void)frow_l;
            (void)grow_l;
            (void)claim_l;
            const int p = __float_as_int(rec.z);
            const float b0 = rec.x, b1 = rec.y;
            const float tg0 = rtg.x, tg1 = rtg.y, tg2 = rtg.z;
            // the next tile's records are in flight while this one shades
            const bool nv = end + j < r1;
            const float4 next = nv ? recs4[2 * (end + j)] : make_float4(0.f, 0.f, 0.f, 0.f);
            const float4 ntg = g == 0 && nv ? recs4[2 * (end + j) + 1] : make_float4(0.f, 0.f, 0.f, 0.f);
            do {
                // ---- the view direction's SH2 values of pixel j (the inputs after the features)
                float shv[9];
                if constexpr (K::SH) {
                    const Views& vw = a.views;
                    int b = 0, base = 0, Wb = vw.W[0], Hb = vw.H[0];
                    for (int k = 1; k < vw.n; ++k)
                        if (p >= vw.pix_base[k]) { b = k; base = vw.pix_base[k]; Wb = vw.W[k]; Hb = vw.H[k]; }
                    const int local = p - base, py = local / Wb, px = local - py * Wb;
                    sh2_basis(a.inv_mvp + 16 * b, pixel_ndc(px, Wb), pixel_ndc(py, Hb), shv);
                }

                // ---- gather from the LDS row: lane (g, j) holds inputs 16 tt + 4 g + r of pixel j
                f32x4 x[NI];
                Corners crs[NI];
                for (int tt = 0; tt < NI; ++tt) {
                    x[tt] = f32x4{0.f, 0.f, 0.f, 0.f};
                    crs[tt] = Corners{{0u, 0u, 0u}, {0.f, 0.f, 0.f}};
                    const int f0 = 16 * tt + 4 * g;
                    if (valid && f0 < DF) {
                        const int lv = f0 / C, c = f0 % C;
                        const Corners cr = lattice_corners(b0, b1, static_cast<uint32_t>(a.lo + lv));
                        crs[tt] = cr;
                        float4 v0, v1, v2;
                        if (FB && f0 == 0) {
                            // channels 0-3 of the first level feed the residual
                            // blend, i.e. the output colour directly: with bf16
                            // rows they are read from the fp32 master (three
                            // 16-byte loads per pixel, cache-resident within a
                            // triangle), so bf16 storage does not reach the colour
                            const float* fg = a.feat + (long long)tl * row;
                            v0 = *reinterpret_cast<const float4*>(fg + cr.idx[0] * C);
                            v1 = *reinterpret_cast<const float4*>(fg + cr.idx[1] * C);
                            v2 = *reinterpret_cast<const float4*>(fg + cr.idx[2] * C);
                        } else if constexpr (DIRECT) {
                            const float* fg = a.feat + (long long)tA * row + loff[lv] * C + c;
                            v0 = ld4_global<C>(fg + cr.idx[0] * C, c);
                            v1 = ld4_global<C>(fg + cr.idx[1] * C, c);
                            v2 = ld4_global<C>(fg + cr.idx[2] * C, c);
                        } else {
                            const FT* fl = frow_l + loff[lv] * CP + c;
                            v0 = ld4(fl + cr.idx[0] * CP);
                            v1 = ld4(fl + cr.idx[1] * CP);
                            v2 = ld4(fl + cr.idx[2] * CP);
                        }
                        float e0 = 0.f, e1 = 0.f, e2 = 0.f, e3 = 0.f;
                        e0 += cr.w[0] * v0.x; e0 += cr.w[1] * v1.x; e0 += cr.w[2] * v2.x;
                        e1 += cr.w[0] * v0.y; e1 += cr.w[1] * v1.y; e1 += cr.w[2] * v2.y;
                        e2 += cr.w[0] * v0.z; e2 += cr.w[1] * v1.z; e2 += cr.w[2] * v2.z;
                        e3 += cr.w[0] * v0.w; e3 += cr.w[1] * v1.w; e3 += cr.w[2] * v2.w;
                        x[tt] = f32x4{e0, e1, e2, e3};
                    }
                    if constexpr (K::SH) {
                        for (int r = 0; r < 4; ++r) {
                            const int k = f0 + r - DF;
                            if (valid && k >= 0 && k < 9) x[tt][r] = pick9(shv, k);
                        }
                    }
                }

                // ---- forward
                f32x4 h1[NH], h2[NH], y[1];
                for (int m = 0; m < NH; ++m)
                    for (int r = 0; r < 4; ++r) { h1[m][r] = a.b1[16 * m + 4 * g + r]; h2[m][r] = a.b2[16 * m + 4 * g + r]; }
                for (int r = 0; r < 4; ++r) y[0][r] = (g == 0 && r < 3) ? a.b3[r] : 0.f;
                dense<K, NH, NI>(blobs, K::L1, lane, x, h1);
                for (int m = 0; m < NH; ++m)
                    for (int r = 0; r < 4; ++r) h1[m][r] = fmaxf(h1[m][r], 0.f);
                dense<K, NH, NH>(blobs, K::L2, lane, h1, h2);
                for (int m = 0; m < NH; ++m)
                    for (int r = 0; r < 4; ++r) h2[m][r] = fmaxf(h2[m][r], 0.f);
                dense<K, 1, NH>(blobs, K::L3, lane, h2, y);

                // ---- blend, loss and output gradient (lanes of group 0 own pixel j)
                f32x4 d3[1] = {f32x4{0.f, 0.f, 0.f, 0.f}};
                float dxr[4] = {0.f, 0.f, 0.f, 0.f};
                if (g == 0 && valid) {
                    const float res = x[0][3];
                    const float tg[3] = {tg0, tg1, tg2};
                    float* col = a.color + (long long)p * 3;
                    float* gc = a.grad_color + (long long)p * 3;
                    float l1 = 0.f, dres = 0.f;
                    for (int c = 0; c < 3; ++c) {
                        const float sg = sigmoidf(y[0][c]);
                        const float rgb = x[0][c];
                        const float out = (1.0f - res) * rgb + res * sg;
                        float go;
                        if (SPLIT && a.ext_grad) {
                            go = a.ext_grad[(long long)p * 3 + c];
                        } else {
                            const float d = out - tg[c];
                            loss += d * d;
                            l1 += fabsf(d);
                            go = a.grad_scale * d;
                            col[c] = out;
                            gc[c] = go;
                        }
                        dxr[c] = go * (1.0f - res);
                        dres += go * (sg - rgb);
                        d3[0][c] = go * res * sg * (1.0f - sg);
                    }
                    dxr[3] = dres;
                    if (!(SPLIT && a.ext_grad)) a.pix_loss[p] = l1;
                    if constexpr (VA_ON) {
                        if (VA) {
                            // the pixel's own triangle's opacity term (src/cuda/rasterize.cu:923-971
                            // for the kept fragment), added into its slot's LDS gradient row
                            const Corners ca = lattice_corners(b0, b1, static_cast<uint32_t>(a.alpha_level));
                            float* av = sl ? va1 : va0;
                            float al = 0.f;
                            al += ca.w[0] * av[ca.idx[0]];
                            al += ca.w[1] * av[ca.idx[1]];
                            al += ca.w[2] * av[ca.idx[2]];
                            const float ga = l1 / (al > 1e-5f ? al : 1e-5f);
                            atomicAdd(av + Sa + ca.idx[0], ga * ca.w[0]);
                            atomicAdd(av + Sa + ca.idx[1], ga * ca.w[1]);
                            atomicAdd(av + Sa + ca.idx[2], ga * ca.w[2]);
                        }
                    }
                }
                if (SPLIT && a.fwd_only) break;   // wave-uniform: colour only, no adjoint

                // ---- backward through the MLP
                f32x4 dh2[NH], dh1[NH], dx[NI];
                for (int m = 0; m < NH; ++m) dh2[m] = dh1[m] = f32x4{0.f, 0.f, 0.f, 0.f};
                for (int m = 0; m < NI; ++m) dx[m] = f32x4{0.f, 0.f, 0.f, 0.f};
                dense<K, NH, 1>(blobs, K::L3T, lane, d3, dh2);
                for (int m = 0; m < NH; ++m)
                    for (int r = 0; r < 4; ++r) dh2[m][r] = h2[m][r] > 0.f ? dh2[m][r] : 0.f;
                dense<K, NH, NH>(blobs, K::L2T, lane, dh2, dh1);
                for (int m = 0; m < NH; ++m)
                    for (int r = 0; r < 4; ++r) dh1[m][r] = h1[m][r] > 0.f ? dh1[m][r] : 0.f;
                dense<K, NI, NH>(blobs, K::L1T, lane, dh1, dx);
                if (g == 0)
                    for (int r = 0; r < 4; ++r) dx[0][r] += dxr[r];

                // ---- scatter the input gradient onto the triangle's gradient row.
                // Lane (g, j) holds dX of pixel j for 4 channels of one level and
                // that level's lattice corners.  The lattices of levels 0-2 (3,
                // 6 and 15 samples) are shared by most pixels of a tile: there
                // the scatter is a matrix product, below.  Finer levels add into
                // the LDS row.  LDS float atomics are slow there (the LDS stays
                // busy ~2 cycles per lane), so each lane first claims its
                // 4-channel block (it writes its lane id into the block's claim
                // word and reads it back); the one lane that reads back its own
                // id owns the block and adds with a plain 16-byte
                // read-modify-write; the losers claim again (CLAIM_ROUNDS) and
                // what is left adds with atomics.
    #pragma unroll
                for (int tt = 0; tt < NI; ++tt) {
                    const int f0 = 16 * tt + 4 * g;
                    const bool in = valid && f0 < DF;
                    const int lv = in ? f0 / C : 0, c = f0 % C;
                    const Corners& cr = crs[tt];
                    if constexpr (K::SH || C != CP) {
                        // only feature inputs scatter (not the SH2 values; no padding channel)
                        for (int r = 0; r < 4; ++r)
                            if (f0 + r >= DF || c + r >= C) dx[tt][r] = 0.f;
                    }
                    if constexpr (DIRECT) {
                        if (in) {
                            float* gg = a.grad_feat + (long long)tA * row + loff[lv] * C + c;
                            for (int k = 0; k < 3; ++k)
                                for (int r = 0; r < 4; ++r)
                                    if (c + r < C) unsafeAtomicAdd(gg + cr.idx[k] * C + r, cr.w[k] * dx[tt][r]);
                        }
                        continue;
                    }
                    bool fine;
                    if constexpr (MF) {
                        fine = in && a.lo + lv > kMfmaTopLevel;
                    } else {
                        // a pixel of the next triangle (a tile crossing into it)
                        // scatters its coarse levels like the fine ones
                        const bool coarse = tt == 0 && a.lo + lv <= 1 && inA;
                        if (in && coarse) {
                            for (int vtx = 0; vtx < 6; ++vtx) {
                                const float wv = (cr.idx[0] == (uint32_t)vtx ? cr.w[0] : 0.f) +
                                                 (cr.idx[1] == (uint32_t)vtx ? cr.w[1] : 0.f) +
                                                 (cr.idx[2] == (uint32_t)vtx ? cr.w[2] : 0.f);
                                for (int r = 0; r < 4; ++r) racc[vtx][r] += wv * dx[0][r];
                            }
                        }
                        fine = in && !coarse;
                    }
                    if constexpr (!P::CLAIM) {
                        // no room for the claim table: LDS atomics
                        if (fine)
                            for (int k = 0; k < 3; ++k) {
                                float* gl = grow_l + grad_slot<CP>(loff[lv] + cr.idx[k], c);
                                for (int r = 0; r < 4; ++r) atomicAdd(gl + r, cr.w[k] * dx[tt][r]);
                            }
                        continue;
                    }
                    if (__ballot(fine) == 0ull) continue;   // wave-uniform
                    int slot[3];
                    for (int k = 0; k < 3; ++k) slot[k] = fine ? grad_slot<CP>(loff[lv] + cr.idx[k], c) : 0;
                    // each round: every pending lane claims its blocks, the owners
                    // add and are done, the losers try again in the next round
                    bool pend[3] = {fine, fine, fine};
    #pragma unroll
                    for (int round = 0; round < K::CLAIM_ROUNDS; ++round) {
                        if (round > 0 && __ballot(pend[0] || pend[1] || pend[2]) == 0ull) break;   // wave-uniform
                        for (int k = 0; k < 3; ++k)
                            if (pend[k])
                                __hip_atomic_store(claim_l + (slot[k] >> 2), lane, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_WAVEFRONT);
                        wave_lds_sync();
                        bool own[3];
                        for (int k = 0; k < 3; ++k)
                            own[k] = pend[k] && __hip_atomic_load(claim_l + (slot[k] >> 2), __ATOMIC_RELAXED,
                                                                  __HIP_MEMORY_SCOPE_WAVEFRONT) == lane;
                        for (int k = 0; k < 3; ++k)
                            if (own[k]) {
                                float4* q = reinterpret_cast<float4*>(grow_l + slot[k]);
                                float4 v = *q;
                                v.x += cr.w[k] * dx[tt][0];
                                v.y += cr.w[k] * dx[tt][1];
                                v.z += cr.w[k] * dx[tt][2];
                                v.w += cr.w[k] * dx[tt][3];
                                *q = v;
                                pend[k] = false;
                            }
                        wave_lds_sync();
                    }
                    for (int k = 0; k < 3; ++k)
                        if (pend[k])
                            for (int r = 0; r < 4; ++r) atomicAdd(grow_l + slot[k] + r, cr.w[k] * dx[tt][r]);
                }

                // ---- levels 0-2: G[s][f] += sum_p W[s][p] dX[p][f] on the fp32
                // MFMA, samples s on M, the level's input tile's 16 feature
                // columns f on N, the tile's pixels p on K (four steps of four).
                // W[s][p] is pixel p's barycentric weight at sample s (the sum
                // over its corners that are s).  dX passes through the stage
                // region transposed ([pixel][f], rows of 20 floats) and the
                // corners through a per-pixel table, so that lane (g, i) of step
                // st can form W[i][4 st + g] and read dX[4 st + g][i].  Columns
                // of other levels sharing the tile accumulate unused values.
                const int nmf = kMfmaTopLevel + 1 - a.lo;   // levels scattered by MFMA (wave-uniform)
                if constexpr (!DIRECT && MF) {
                    float* trn = reinterpret_cast<float*>(stage);
                    int4* cwi = reinterpret_cast<int4*>(trn + 16 * 20);
                    float4* cwf = reinterpret_cast<float4*>(cwi + 16);
    #pragma unroll
                    for (int q = 0; q < NMF; ++q) {
                        if (q >= nmf) break;
                        const int tq = q * C / 16, c0 = q * C - 16 * tq;
                        if (q == 0 || tq != (q - 1) * C / 16)
                            *reinterpret_cast<f32x4*>(trn + j * 20 + 4 * g) = dx[tq];
                        if (4 * g == (c0 & ~3)) {
                            const Corners& cr = crs[tq];
                            cwi[j] = make_int4(static_cast<int>(cr.idx[0]), static_cast<int>(cr.idx[1]),
                                               static_cast<int>(cr.idx[2]), 0);
                            cwf[j] = make_float4(cr.w[0], cr.w[1], cr.w[2], 0.f);
                        }
                        wave_lds_sync();
    #pragma unroll
                        for (int st = 0; st < 4; ++st) {
                            const int p = 4 * st + g;
                            const int4 ci = cwi[p];
                            const float4 cf = cwf[p];
                            const float w = (ci.x == j ? cf.x : 0.f) + (ci.y == j ? cf.y : 0.f) + (ci.z == j ? cf.z : 0.f);
                            sacc[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(w, trn[p * 20 + j], sacc[q], 0, 0, 0);
                        }
                        wave_lds_sync();
                    }
                }
                if constexpr (!MF) {
                    // bias gradients per lane until the end
                    for (int m = 0; m < NH; ++m) { gb1[m] += dh1[m]; gb2[m] += dh2[m]; }
                    gb3 += d3[0];
                }

                // ---- weight gradients: transpose (delta, activation) pairs through
                // LDS one at a time and contract over the 16 pixels
                {
                    constexpr int DINP = K::DINP;
                    ST* sD = stage;
                    ST* sA = stage + K::region(DH);
                    for (int m = 0; m < NH; ++m) stage_put<K>(sD, DH, 16 * m + 4 * g, j, dh1[m]);
                    for (int tt = 0; tt < NI; ++tt) stage_put<K>(sA, DINP, 16 * tt + 4 * g, j, x[tt]);
                    wave_lds_sync();
                    for (int m = 0; m < NH; ++m) {
                        for (int nn = 0; nn < NI; ++nn)
                            gW1[m][nn] = stage_outer<K>(sD, DH, 16 * m, sA, DINP, 16 * nn, lane, gW1[m][nn]);
                        if constexpr (MF) gbias = stage_sum<K>(sD, DH, 16 * m, m, lane, gbias);
                    }
                    wave_lds_sync();
                    for (int m = 0; m < NH; ++m) {
                        stage_put<K>(sD, DH, 16 * m + 4 * g, j, dh2[m]);
                        stage_put<K>(sA, DH, 16 * m + 4 * g, j, h1[m]);
                    }
                    wave_lds_sync();
                    for (int m = 0; m < NH; ++m) {
                        for (int nn = 0; nn < NH; ++nn)
                            gW2[m][nn] = stage_outer<K>(sD, DH, 16 * m, sA, DH, 16 * nn, lane, gW2[m][nn]);
                        if constexpr (MF) gbias = stage_sum<K>(sD, DH, 16 * m, NH + m, lane, gbias);
                    }
                    wave_lds_sync();
                    ST* sD3 = stage;
                    ST* sH2 = stage + K::region(16);
                    stage_put<K>(sD3, 16, 4 * g, j, d3[0]);
                    for (int m = 0; m < NH; ++m) stage_put<K>(sH2, DH, 16 * m + 4 * g, j, h2[m]);
                    wave_lds_sync();
                    for (int m = 0; m < NH; ++m)
                        gW3[m] = stage_outer<K>(sD3, 16, 0, sH2, DH, 16 * m, lane, gW3[m]);
                    if constexpr (MF) gbias = stage_sum<K>(sD3, 16, 0, 2 * NH, lane, gbias);
                    wave_lds_sync();
                }
            } while (false);
            rec = next;
            rtg = ntg;
        }
        pos = end;

        // ---- the triangles this tile completed: A, and (a tile that ran
        // through all of N) N after it; at the range's end A is done
        // whether or not its pixels continue in the next wave's range
        while (pos >= min(teA, r1)) {
            if (!(SPLIT && a.fwd_only) && !DIRECT) {
                float* grow = sA ? grow1 : grow0;
                // ---- the coarse levels' blocks into the LDS row
                if constexpr (MF) {
                    add_mf_blocks<K, NMF>(sacc, kMfmaTopLevel + 1 - a.lo, a.lo, loff, grow, g, j);
                    for (int q = 0; q < NMF; ++q) sacc[q] = f32x4{0.f, 0.f, 0.f, 0.f};
                    wave_lds_sync();
                } else {
                    // the two coarsest levels' sums: reduce over the 16 pixel
                    // lanes of each lane group, then one lane per group adds
                    // them to the LDS row
                    const int f0 = 4 * g;
                    const int lv = f0 / C, c = f0 % C;
                    const bool coarse = f0 < DF && a.lo + lv <= 1;
                    const int S = coarse ? static_cast<int>(feats_at_level(a.lo + lv)) : 0;
                    for (int vtx = 0; vtx < 6; ++vtx)
                        for (int r = 0; r < 4; ++r) {
                            racc[vtx][r] = row16_sum(racc[vtx][r]);
                        }
                    wave_lds_sync();
                    if (coarse && j == 0) {
                        for (int vtx = 0; vtx < 6; ++vtx)
                            if (vtx < S) {
                                float* gl = grow + grad_slot<CP>(loff[lv] + vtx, c);
                                for (int r = 0; r < 4; ++r) gl[r] += racc[vtx][r];
                            }
                    }
                    for (int vtx = 0; vtx < 6; ++vtx)
                        for (int r = 0; r < 4; ++r) racc[vtx][r] = 0.f;
                    wave_lds_sync();
                }

                // ---- PF: the next triangle's bf16 row goes out before this
                // triangle's gradient row: loads issued ahead of the stores
                // complete without waiting for them (the memory counter is in
                // order), so the row's latency overlaps the stores and the row
                // clearing
                if constexpr (PF) {
                    if (teA < r1 && row / 4 <= PF_N * kWave) {
                        const uint2* src = reinterpret_cast<const uint2*>(a.feat_bf16 + (long long)tN * row);
#pragma unroll
                        for (int u = 0; u < PF_N; ++u)
                            pf[u] = lane + u * kWave < row / 4 ? src[lane + u * kWave] : make_uint2(0u, 0u);
                        have_pf = true;
                    }
                }

                // ---- the triangle's feature gradient row
                if (ownA && C == CP) {
                    float4* dst = reinterpret_cast<float4*>(a.grad_feat + (long long)tA * row);
                    for (int i = lane; i < row / 4; i += kWave)
                        dst[i] = *reinterpret_cast<const float4*>(grow + grad_slot<CP>(4 * i / C, 4 * i % C));
                } else {
                    float* dst = a.grad_feat + (long long)tA * row;
                    for (int i = lane; i < row; i += kWave) {
                        const int sm = i / C, c = i - sm * C;
                        const float gv = grow[grad_slot<CP>(sm, c & ~3) + (c & 3)];
                        if (ownA) dst[i] = gv;
                        else if (gv != 0.f) unsafeAtomicAdd(&dst[i], gv);
                    }
                }
                // ---- the triangle's visible-fragment opacity gradient row
                // (the table is zeroed before the step; the opacity-gradient
                // pass adds the other triangles' fragments after this kernel)
                if constexpr (VA_ON) {
                    if (VA) {
                        const float* ag = (sA ? va1 : va0) + Sa;
                        float* dst = a.grad_alpha + (long long)tA * Sa;
                        for (int i = lane; i < Sa; i += kWave) {
                            const float gv = ag[i];
                            if (ownA) dst[i] = gv;
                            else if (gv != 0.f) unsafeAtomicAdd(&dst[i], gv);
                        }
                    }
                }
                wave_lds_sync();
            }
            // at the range's end, a staged N (its first pixels were in the
            // last tile) still has its row to add
            if (pos >= r1 && !(XT && haveN)) break;
            // ---- N becomes A
            tA = tN;
            teA = teN;
            ownA = teA <= r1;
            if (XT && haveN) {
                sA = 1 - sA;
                haveN = false;
            } else {
                begin_row(sA, tA);
            }
            if (teA < r1) {
                tN = a.list[teA].tri;
                teN = a.list[teA].seg_end;
            }
        }
    }

    if constexpr (!MF) {
        // reduce the per-lane bias gradients over the 16 pixels of each lane group
        for (int m = 0; m < NH; ++m)
            for (int r = 0; r < 4; ++r) {
                gb1[m][r] = row16_sum(gb1[m][r]);
                gb2[m][r] = row16_sum(gb2[m][r]);
            }
        for (int r = 0; r < 4; ++r) gb3[r] = row16_sum(gb3[r]);
    }
    for (int d = 32; d > 0; d >>= 1) loss += __shfl_xor(loss, d, kWave);

    // ---- workgroup reduction, one partial per block: every wave stores its
    // accumulators into its own LDS copy of the parameter vector (each entry
    // is held by exactly one lane, so plain stores; the whole LDS is free by
    // now), then the block sums the four copies
    __syncthreads();
    float* red = smem + wave * K::NPARAM;
    float* rW1 = red;
    float* rW2 = rW1 + DH * DIN;
    float* rW3 = rW2 + DH * DH;
    float* rb1 = rW3 + 3 * DH;
    float* rb2 = rb1 + DH;
    float* rb3 = rb2 + DH;
    float* rl = rb3 + 3;
    for (int m = 0; m < NH; ++m) {
        for (int r = 0; r < 4; ++r) {
            const int u = 16 * m + 4 * g + r;   // accumulator row
            for (int nn = 0; nn < NI; ++nn) {
                const int f = 16 * nn + j;
                if (f < DIN) rW1[u * DIN + f] = gW1[m][nn][r];
            }
            for (int nn = 0; nn < NH; ++nn) rW2[u * DH + 16 * nn + j] = gW2[m][nn][r];
            if (!MF && j == 0) {
                rb1[u] = gb1[m][r];
                rb2[u] = gb2[m][r];
            }
        }
        // gW3 rows are output channels 4 g + r (only 0..2 exist), columns hidden units
        if (g == 0)
            for (int r = 0; r < 3; ++r) rW3[r * DH + 16 * m + j] = gW3[m][r];
    }
    if constexpr (MF) {
        // bias block: lane (g, n) holds rows 4 g + r of column n
        for (int r = 0; r < 4; ++r) {
            if (j < NH) rb1[16 * j + 4 * g + r] = gbias[r];
            else if (j < 2 * NH) rb2[16 * (j - NH) + 4 * g + r] = gbias[r];
            else if (j == 2 * NH && g == 0 && r < 3) rb3[r] = gbias[r];
        }
    } else if (g == 0 && j == 0) {
        for (int r = 0; r < 3; ++r) rb3[r] = gb3[r];
    }
    if (lane == 0) *rl = loss;
    __syncthreads();
    float* out = a.partials + (long long)blockIdx.x * K::NPARAM;
    for (int i = threadIdx.x; i < K::NPARAM; i += blockDim.x) {
        float v = smem[i];
        for (int w = 1; w < kWavesPerBlock; ++w) v += smem[w * K::NPARAM + i];
        out[i] = v;
    }
}

template <class K, bool FB, bool SPLIT, bool DIRECT>
hipError_t shade_launch(const ShadeArgs& a, int nblocks, int rowp, int ns, size_t bytes, hipStream_t s) {
    static size_t opted_in = 65536;   // per instance: more LDS needs the kernel's opt-in
    if (bytes > opted_in) {
        const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k_shade<K, FB, SPLIT, DIRECT>),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                                 static_cast<int>(bytes));
        if (e != hipSuccess) return e;
        opted_in = bytes;
    }
    k_shade<K, FB, SPLIT, DIRECT><<<nblocks, kWave * kWavesPerBlock, bytes, s>>>(a, rowp, ns);
    return hipGetLastError();
}

template <class K, bool FB, bool DIRECT>
hipError_t shade_forms(const ShadeArgs& a, int nblocks, int rowp, int ns, size_t bytes, hipStream_t s) {
    if (a.fwd_only || a.ext_grad) return shade_launch<K, FB, true, DIRECT>(a, nblocks, rowp, ns, bytes, s);
    return shade_launch<K, FB, false, DIRECT>(a, nblocks, rowp, ns, bytes, s);
}

// Staged rows when the LDS plan fits; single-level rows that do not fit are
// gathered and scattered in global memory instead (DIRECT).  The kernels
// without MF take a second row slot (tiles running on into the next
// triangle) when it keeps their workgroups per CU.
//
// The visible-opacity rows (the fused single pass with a.alpha set) are
// taken when they fit beside the chosen slots: two slots with them, two
// without, one with, one without.
template <class K, bool FB>
hipError_t shade_dispatch(const ShadeArgs& a0, int nblocks, hipStream_t s, bool* va_out) {
    using P = Plan<K, FB>;
    const int rowp = (a0.S_tot * K::CP + 3) & ~3;
    const size_t red = sizeof(float) * kWavesPerBlock * K::NPARAM;
    // (the kernel compiles the rows for the two-waves-per-SIMD forms only)
    const bool want_va = !K::MF && a0.alpha && a0.grad_alpha && !a0.fwd_only && !a0.ext_grad && a0.Sa > 0;
    const auto lds_of = [&](int rp, int ns, bool va) {
        return std::max(red, P::BLOB_BYTES + static_cast<size_t>(kWavesPerBlock) *
                                                 (P::wave_bytes(rp, ns) + (va ? 4 * ns * va_floats(a0.Sa) : 0)));
    };
    ShadeArgs a = a0;
    const auto go = [&](int ns, bool va) {
        if (!va) a.alpha = nullptr;
        if (va_out) *va_out = va;
        return shade_forms<K, FB, false>(a, nblocks, rowp, ns, lds_of(rowp, ns, va), s);
    };
    if (P::NS == 2) {
        const size_t lim = 160u * 1024u / K::MIN_BLOCKS;
        if (want_va && lds_of(rowp, 2, true) <= lim) return go(2, true);
        if (lds_of(rowp, 2, false) <= lim) return go(2, false);
    }
    if (want_va && lds_of(rowp, 1, true) <= 160 * 1024) return go(1, true);
    if (lds_of(rowp, 1, false) <= 160 * 1024) return go(1, false);
    if constexpr (K::NL == 1 && !FB) {
        if (lds_of(0, 1, false) <= 160 * 1024) {
            a.alpha = nullptr;
            if (va_out) *va_out = false;
            return shade_forms<K, FB, true>(a, nblocks, 0, 1, lds_of(0, 1, false), s);
        }
    }
    return hipErrorInvalidValue;
}

// One configuration: fp32 MLP, bf16 MLP on fp32 rows, or bf16 MLP on bf16
// rows (whole 4-channel blocks only).
template <int NL, int C, int DH, bool SH>
hipError_t shade_config(bool bf16, const ShadeArgs& a, int nblocks, hipStream_t s, bool* va) {
    if (!bf16) return shade_dispatch<Cfg<NL, C, DH, false, SH>, false>(a, nblocks, s, va);
    if (a.feat_bf16) {
        if constexpr (C % 4 == 0) return shade_dispatch<Cfg<NL, C, DH, true, SH>, true>(a, nblocks, s, va);
        return hipErrorInvalidValue;
    }
    return shade_dispatch<Cfg<NL, C, DH, true, SH>, false>(a, nblocks, s, va);
}

// The compiled configurations: (levels, channels per level, hidden width,
// SH2 inputs after the features).  (1, 7, 16, SH2) is the reference's own
// network: 7 accumulated feature channels + 9 SH2 values into 16 -> 16 -> 16
// -> 3 (examples/02_synthetic.py:290,305).
#define DS_SHADE_CONFIGS(X) \
    X(1, 8, 16, false)      \
    X(4, 8, 32, false)      \
    X(1, 8, 32, false)      \
    X(4, 4, 32, false)      \
    X(2, 8, 32, false)      \
    X(5, 16, 64, false)     \
    X(1, 7, 16, true)

}  // namespace

bool shade_supported(int nl, int c, int dh, bool sh) {
#define X(NL, C, DH, SH) if (nl == NL && c == C && dh == DH && sh == SH) return true;
    DS_SHADE_CONFIGS(X)
#undef X
    return false;
}

size_t shade_blob_floats(int nl, int c, int dh, bool sh) {
#define X(NL, C, DH, SH) \
    if (nl == NL && c == C && dh == DH && sh == SH) return Cfg<NL, C, DH, false, SH>::BLOB_ELEMS;
    DS_SHADE_CONFIGS(X)
#undef X
    return 0;
}

int shade_blocks_per_cu(int nl, int c, int dh, bool sh) {
#define X(NL, C, DH, SH) \
    if (nl == NL && c == C && dh == DH && sh == SH) return Cfg<NL, C, DH, false, SH>::MIN_BLOCKS;
    DS_SHADE_CONFIGS(X)
#undef X
    return 1;
}

long long shade_seg_cost(int nl, int c, int dh, bool sh) {
#define X(NL, C, DH, SH) \
    if (nl == NL && c == C && dh == DH && sh == SH) return Cfg<NL, C, DH, false, SH>::SEG_COST;
    DS_SHADE_CONFIGS(X)
#undef X
    return 16;
}

int shade_nparam(int nl, int c, int dh, bool sh) {
    const int din = nl * c + (sh ? 9 : 0);
    return dh * din + dh * dh + 3 * dh + dh + dh + 3 + 1;
}

namespace {
template <class K, bool FB>
bool direct_rows(int S_tot) {
    using P = Plan<K, FB>;
    const int rowp = (S_tot * K::CP + 3) & ~3;
    const size_t red = sizeof(float) * kWavesPerBlock * K::NPARAM;
    const size_t lds = P::BLOB_BYTES + static_cast<size_t>(kWavesPerBlock) * P::wave_bytes(rowp, 1);
    return std::max(lds, red) > 160 * 1024;
}
template <int NL, int C, int DH, bool SH>
bool config_direct(bool bf16, bool feat_bf16, int S_tot) {
    if (!bf16) return direct_rows<Cfg<NL, C, DH, false, SH>, false>(S_tot);
    if (feat_bf16 && C % 4 == 0) return direct_rows<Cfg<NL, C, DH, true, SH>, true>(S_tot);
    return direct_rows<Cfg<NL, C, DH, true, SH>, false>(S_tot);
}
}  // namespace

bool shade_direct_rows(int nl, int c, int dh, bool sh, bool bf16, bool feat_bf16, int S_tot) {
#define X(NL, C, DH, SH) \
    if (nl == NL && c == C && dh == DH && sh == SH) return config_direct<NL, C, DH, SH>(bf16, feat_bf16, S_tot);
    DS_SHADE_CONFIGS(X)
#undef X
    return false;
}

hipError_t launch_shade(int nl, int c, int dh, bool sh, bool bf16, const ShadeArgs& a, int nblocks,
                        hipStream_t s, bool* visible_alpha) {
    if (visible_alpha) *visible_alpha = false;
#define X(NL, C, DH, SH) \
    if (nl == NL && c == C && dh == DH && sh == SH) \
        return shade_config<NL, C, DH, SH>(bf16, a, nblocks, s, visible_alpha);
    DS_SHADE_CONFIGS(X)
#undef X
    return hipErrorInvalidValue;
}


}  // namespace ds
