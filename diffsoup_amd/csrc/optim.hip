// Fused optimiser step over every parameter tensor of a training step:
// torch.optim.Adam's update for element-wise groups and VectorAdam's
// (second moment pooled over the last dimension) for vector groups, in one
// launch, with the step counter kept in device memory so that a captured
// graph advances it on every replay.
//
// Reference: python/diffsoup/optimize.py:33-68 (VectorAdam), and the Adam
// groups of examples/02_synthetic.py:310-315 (torch.optim.Adam defaults:
// betas (0.9, 0.999), eps 1e-8, no weight decay).
#include <hip/hip_runtime.h>
#include <math.h>

#include <algorithm>

#include "diffsoup_hip.h"

namespace ds {
namespace {

constexpr int kMaxGroups = 16;
constexpr int kThreads = 256;
// Three launch shapes by parameter count (measured over 9.1M, 28M, 93M,
// 355M and 1.07G parameters): up to about 16M parameters one float4 per
// thread per chunk and a grid of 16 blocks per CU (46 us vs 49 us at 8 per
// CU); up to 128M, two float4s per thread, issued together, and 4 blocks per
// CU (522 us vs 585 us at 93M; 186 vs 211 us at 28M against four float4s);
// above, four float4s per thread and 2 blocks per CU (355M: 2.26 -> 2.14 ms,
// 1.07G: 6.73 -> 6.44 ms).
constexpr int kChunkSmall = 4 * 1 * kThreads, kChunkLarge = 4 * 2 * kThreads, kChunkHuge = 4 * 4 * kThreads;
constexpr int kBlocksSmall = 256 * 16, kBlocksLarge = 256 * 4, kBlocksHuge = 256 * 2;
constexpr long long kLargeUnits = 16ll << 20, kHugeUnits = 128ll << 20;

struct Group {
    const float* p;
    float* out;       // updated parameter (== p for an in-place update)
    unsigned short* out16;   // optional bf16 copy of the updated parameter
    const float* g;
    float* m;
    float* v;
    long long units;  // elements (Adam) or vectors (VectorAdam)
    int dim;          // 0: Adam; k > 0: VectorAdam over vectors of k floats
    double lr, b1, b2;     // bias corrections are formed in double
    float b1f, b2f, w1, w2, eps;  // fp32 forms: beta, 1 - beta (rounded from double), eps
    const float* lr_mult;  // optional device multiplier of lr
    long long decay_steps; // > 0: exponential decay of lr with the iteration
    double decay_final;
    long long block_base;  // first chunk of this group
};

struct Table {
    Group grp[kMaxGroups];
    int n;
};

typedef float f4v __attribute__((ext_vector_type(4)));
typedef unsigned int u2v __attribute__((ext_vector_type(2)));

// The updated parameters and moments are not read again before the next
// step's optimiser launch, so their stores carry the non-temporal hint and
// leave the caches to the next step's raster and shading kernels (measured:
// the step gains ~0.7%; non-temporal loads as well made the update itself
// slower).
__device__ __forceinline__ void st_nt(float* p, float4 v) {
    f4v x = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(x, reinterpret_cast<f4v*>(p));
}
__device__ __forceinline__ void st_nt2(unsigned short* p, uint2 v) {
    u2v x = {v.x, v.y};
    __builtin_nontemporal_store(x, reinterpret_cast<u2v*>(p));
}

// bf16 bits of x, rounded to nearest even (NaN stays NaN).
__device__ __forceinline__ uint32_t bf16_bits(float x) {
    const uint32_t u = __float_as_uint(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (u >> 16) | 0x40u;
    return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
}

__device__ __forceinline__ float adam_update(const Group& G, float step_size, float bc2_sqrt, float p, float g,
                                             float& m, float& v) {
    // torch.optim.Adam (single-tensor form): m.lerp_(g, 1 - b1);
    // v = v * b2 + (1 - b2) g g; p -= (lr / bc1) * m / (sqrt(v) / sqrt(bc2) + eps)
    m = m + G.w1 * (g - m);
    v = v * G.b2f;
    v = v + G.w2 * (g * g);
    const float denom = sqrtf(v) / bc2_sqrt + G.eps;
    return p - step_size * (m / denom);
}

// Persistent over chunks of kChunk units: chunk c belongs to the group
// whose [chunk_base, next chunk_base) holds it.  A block forms the bias
// corrections of every group once (in double, as torch evaluates beta ** step
// on the host), then walks the chunks with a grid stride, so the grid is one
// resident wave of blocks and no tail of stragglers is left.
template <int kPerThread>
__global__ void __launch_bounds__(kThreads) k_adam(Table tab, long long n_chunks,
                                                   const long long* __restrict__ step_before, int advanced) {
    constexpr int kChunk = 4 * kPerThread * kThreads;
    // per group, once per block: the bias corrections (in double, as torch
    // forms beta ** step on the host) and the fp32 factors derived from them
    __shared__ float coef[kMaxGroups][5];   // lr / bc1, sqrt(bc2), bc1, bc2, lr
    if (threadIdx.x < (unsigned)tab.n) {
        const Group& G = tab.grp[threadIdx.x];
        // the iteration: the counter + 1, or the counter itself when an
        // earlier kernel of the stream has already advanced it
        const double t = static_cast<double>(*step_before + (advanced ? 0 : 1));
        const double bc1 = 1.0 - pow(G.b1, t), bc2 = 1.0 - pow(G.b2, t);
        // the learning rate of this iteration: the schedule's multiplier
        // final ^ (i / steps), i = t clamped to [1, steps], then the caller's
        double lr = G.lr;
        if (G.decay_steps > 0) {
            const double n = static_cast<double>(G.decay_steps);
            lr *= pow(G.decay_final, fmin(fmax(t, 1.0), n) / n);
        }
        if (G.lr_mult) lr *= static_cast<double>(*G.lr_mult);
        coef[threadIdx.x][0] = static_cast<float>(lr / bc1);
        coef[threadIdx.x][1] = static_cast<float>(sqrt(bc2));
        coef[threadIdx.x][2] = static_cast<float>(bc1);
        coef[threadIdx.x][3] = static_cast<float>(bc2);
        coef[threadIdx.x][4] = static_cast<float>(lr);
    }
    __syncthreads();
    // kPerThread float4s of each operand per thread, all loaded before any
    // is used: u0 is a multiple of kChunk, so the accesses stay 16-byte aligned
    struct Regs {
        float4 p[kPerThread], g[kPerThread], m[kPerThread], v[kPerThread];
    };
    auto load = [&](const Group& G, long long u0, long long n4, Regs& r) {
#pragma unroll
        for (int k = 0; k < kPerThread; ++k) {
            const long long q = threadIdx.x + k * kThreads;
            if (q < n4) {
                const long long e = u0 + 4 * q;
                r.p[k] = *reinterpret_cast<const float4*>(G.p + e);
                r.g[k] = *reinterpret_cast<const float4*>(G.g + e);
                r.m[k] = *reinterpret_cast<const float4*>(G.m + e);
                r.v[k] = *reinterpret_cast<const float4*>(G.v + e);
            }
        }
    };
    auto update = [&](const Group& G, int gi, long long u0, long long n4, Regs& r) {
        const float step_size = coef[gi][0];
        const float bc2_sqrt = coef[gi][1];
#pragma unroll
        for (int k = 0; k < kPerThread; ++k) {
            const long long q = threadIdx.x + k * kThreads;
            if (q < n4) {
                const long long e = u0 + 4 * q;
                float4 o;
                o.x = adam_update(G, step_size, bc2_sqrt, r.p[k].x, r.g[k].x, r.m[k].x, r.v[k].x);
                o.y = adam_update(G, step_size, bc2_sqrt, r.p[k].y, r.g[k].y, r.m[k].y, r.v[k].y);
                o.z = adam_update(G, step_size, bc2_sqrt, r.p[k].z, r.g[k].z, r.m[k].z, r.v[k].z);
                o.w = adam_update(G, step_size, bc2_sqrt, r.p[k].w, r.g[k].w, r.m[k].w, r.v[k].w);
                st_nt(G.m + e, r.m[k]);
                st_nt(G.v + e, r.v[k]);
                st_nt(G.out + e, o);
                if (G.out16) {
                    uint2 h;
                    h.x = bf16_bits(o.x) | (bf16_bits(o.y) << 16);
                    h.y = bf16_bits(o.z) | (bf16_bits(o.w) << 16);
                    st_nt2(G.out16 + e, h);
                }
            }
        }
    };
    // With two float4s per thread the next chunk of the same group is loaded
    // before this chunk's stores are issued: the memory counter retires in
    // order, so its loads then complete without waiting for them (28M
    // parameters: 185 -> 165 us).  At four float4s per thread the second
    // register set costs more than it hides (355M: 2.13 -> 2.25 ms), and at
    // one the grid's sixteen blocks per CU already hide the stores.
    constexpr bool kPipe = kPerThread == 2;
    Regs cur, nxt;
    bool have = false;   // cur already holds chunk c (loaded during the previous chunk)
    for (long long c = blockIdx.x; c < n_chunks; c += gridDim.x) {
        int gi = 0;
        while (gi + 1 < tab.n && c >= tab.grp[gi + 1].block_base) ++gi;
        const Group& G = tab.grp[gi];
        const long long u0 = (c - G.block_base) * kChunk;
        const long long u1 = min(u0 + kChunk, G.units);
        if (G.dim == 0) {
            const float step_size = coef[gi][0];
            const float bc2_sqrt = coef[gi][1];
            const bool vec = ((reinterpret_cast<uintptr_t>(G.p) | reinterpret_cast<uintptr_t>(G.out) |
                               reinterpret_cast<uintptr_t>(G.g) | reinterpret_cast<uintptr_t>(G.m) |
                               reinterpret_cast<uintptr_t>(G.v)) & 15u) == 0 &&
                             (reinterpret_cast<uintptr_t>(G.out16) & 7u) == 0;
            long long i = u0;
            if (vec) {
                const long long n4 = (u1 - u0) / 4;
                if (!have) load(G, u0, n4, cur);
                have = false;
                if constexpr (kPipe) {
                    const long long cn = c + gridDim.x;
                    if (cn < n_chunks && (gi + 1 >= tab.n || cn < tab.grp[gi + 1].block_base)) {
                        const long long u0n = (cn - G.block_base) * kChunk;
                        load(G, u0n, (min(u0n + kChunk, G.units) - u0n) / 4, nxt);
                        have = true;
                    }
                }
                update(G, gi, u0, n4, cur);
                if constexpr (kPipe) {
                    if (have) cur = nxt;
                }
                i = u0 + 4 * n4;
            }
            for (i += threadIdx.x; i < u1; i += kThreads) {
                float m = G.m[i], v = G.v[i];
                const float o = adam_update(G, step_size, bc2_sqrt, G.p[i], G.g[i], m, v);
                G.m[i] = m;
                G.v[i] = v;
                G.out[i] = o;
                if (G.out16) G.out16[i] = static_cast<unsigned short>(bf16_bits(o));
            }
        } else {
            // VectorAdam: g1 = b1 g1 + (1 - b1) g; g2 = b2 g2 + (1 - b2) |g|^2;
            // p -= lr * (g1 / bc1) / (sqrt(g2 / bc2) + 1e-8)
            const float ibc1 = coef[gi][2], ibc2 = coef[gi][3], lrf = coef[gi][4];
            const int k = G.dim;
            for (long long u = u0 + threadIdx.x; u < u1; u += kThreads) {
                const long long base = u * k;
                float sq = 0.f;
                for (int j = 0; j < k; ++j) {
                    const float g = G.g[base + j];
                    sq += g * g;
                }
                const float g2 = G.v[u] * G.b2f + G.w2 * sq;
                G.v[u] = g2;
                const float den = sqrtf(g2 / ibc2) + G.eps;
                for (int j = 0; j < k; ++j) {
                    const float g = G.g[base + j];
                    const float g1 = G.m[base + j] * G.b1f + G.w1 * g;
                    G.m[base + j] = g1;
                    const float o = G.p[base + j] - lrf * ((g1 / ibc1) / den);
                    G.out[base + j] = o;
                    if (G.out16) G.out16[base + j] = static_cast<unsigned short>(bf16_bits(o));
                }
            }
        }
    }
}

__global__ void k_tick(long long* step) {
    if (threadIdx.x == 0) *step += 1;
}

}  // namespace

int optim_fail(int code, const char* msg);

}  // namespace ds

namespace ds {
namespace {
int adam_step(const ds_adam_group* groups, int n_groups, int64_t* step, void* stream, bool advanced);
}  // namespace
}  // namespace ds

extern "C" int ds_adam_step(const ds_adam_group* groups, int n_groups, int64_t* step, void* stream) {
    return ds::adam_step(groups, n_groups, step, stream, false);
}

extern "C" int ds_adam_step_advanced(const ds_adam_group* groups, int n_groups, int64_t* step, void* stream) {
    return ds::adam_step(groups, n_groups, step, stream, true);
}

namespace ds {
namespace {
int adam_step(const ds_adam_group* groups, int n_groups, int64_t* step, void* stream, bool advanced) {
    if (n_groups < 1 || n_groups > kMaxGroups)
        return optim_fail(DS_ERR_INVALID, "adam: 1..16 parameter groups");
    if (!step) return optim_fail(DS_ERR_INVALID, "adam: step counter is null");
    Table tab{};
    tab.n = n_groups;
    long long units = 0;
    for (int i = 0; i < n_groups; ++i) units += groups[i].n;
    const int shape = units > kHugeUnits ? 2 : units > kLargeUnits ? 1 : 0;
    const int chunk = shape == 2 ? kChunkHuge : shape == 1 ? kChunkLarge : kChunkSmall;
    long long blocks = 0;
    for (int i = 0; i < n_groups; ++i) {
        const ds_adam_group& s = groups[i];
        if (s.n < 0 || s.vector_dim < 0 || (s.vector_dim > 0 && s.n % s.vector_dim) ||
            (s.n > 0 && (!s.param || !s.grad || !s.exp_avg || !s.exp_avg_sq)))
            return optim_fail(DS_ERR_INVALID, "adam: bad group (size, vector width or null pointer)");
        Group& G = tab.grp[i];
        G.p = s.param;
        G.out = s.param_out ? s.param_out : s.param;
        G.out16 = reinterpret_cast<unsigned short*>(s.param_bf16_out);
        G.g = s.grad;
        G.m = s.exp_avg;
        G.v = s.exp_avg_sq;
        G.dim = s.vector_dim;
        G.units = s.vector_dim ? s.n / s.vector_dim : s.n;
        G.lr = s.lr;
        G.b1 = s.beta1;
        G.b2 = s.beta2;
        G.b1f = static_cast<float>(s.beta1);
        G.b2f = static_cast<float>(s.beta2);
        G.w1 = static_cast<float>(1.0 - s.beta1);
        G.w2 = static_cast<float>(1.0 - s.beta2);
        G.eps = static_cast<float>(s.eps);
        G.lr_mult = s.lr_mult;
        G.decay_steps = s.decay_steps;
        G.decay_final = s.decay_final;
        if (s.decay_steps > 0 && !(s.decay_final > 0.0))
            return optim_fail(DS_ERR_INVALID, "adam: decay_final must be > 0 with decay_steps");
        G.block_base = blocks;
        blocks += (G.units + chunk - 1) / chunk;
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    long long* t = reinterpret_cast<long long*>(step);
    if (blocks > 0) {
        if (shape == 2)
            k_adam<4><<<static_cast<unsigned>(std::min<long long>(blocks, kBlocksHuge)), kThreads, 0, st>>>(tab, blocks,
                                                                                                           t, advanced);
        else if (shape == 1)
            k_adam<2><<<static_cast<unsigned>(std::min<long long>(blocks, kBlocksLarge)), kThreads, 0, st>>>(tab, blocks,
                                                                                                            t, advanced);
        else
            k_adam<1><<<static_cast<unsigned>(std::min<long long>(blocks, kBlocksSmall)), kThreads, 0, st>>>(tab, blocks,
                                                                                                            t, advanced);
    }
    if (!advanced) k_tick<<<1, 64, 0, st>>>(t);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? DS_OK : optim_fail(DS_ERR_HIP, hipGetErrorString(e));
}
}  // namespace
}  // namespace ds
