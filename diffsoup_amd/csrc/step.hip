// One fused training step (forward + backward) on the caller's stream:
//
//   k_step_prep     sigmoid of the opacity logits, zero the gradient
//                   accumulators and counters, clip-space projection
//   k_blobs         per-lane MLP weight operands (shade.hip)
//   raster bin      setup / scan / emit (raster.hip)
//   resolve+shade   depth test, background shading, covered-pixel list,
//                   fragment records for the opacity gradient
//   k_shade         gather -> MLP -> blend -> loss gradient, MLP adjoint,
//                   feature scatter
//   edge grad       silhouette / intersection gradients into clip positions
//   opacity aux     score-function opacity gradient over the fragment
//                   records (a second walk of the bins for a tile whose
//                   records overflowed)
//   k_epilogue      MLP weight-gradient reduction, the loss, clip -> world
//                   vertex gradients and the sigmoid backward
//
// No allocation, no host synchronisation: everything lives in the caller's
// workspace and output buffers, so the sequence can be captured in a graph.
// Composition follows examples/02_synthetic.py:354-388 with the features
// concatenated per level; the image loss is MSE (one shading pass that forms
// the loss gradient per pixel) or the reference's 0.8 L1 + 0.2 (1 - SSIM) / 2
// (a colour pass, the loss kernels of ssim.hip over the whole image, then the
// adjoint pass on their gradient).
#include "diffsoup_hip.h"
#include "ds_kernels.h"
#include "ds_math.h"
#include "ds_scan.h"
#include "ds_shade.h"

namespace ds {
namespace {

constexpr int kWave = 64;

// Zero n floats at p (16-byte aligned) from a grid-stride loop: float4 body,
// scalar tail.
__device__ __forceinline__ void zero_floats(float* p, long long n, long long i0, long long stride) {
    if (!p) return;
    float4* p4 = reinterpret_cast<float4*>(p);
    for (long long i = i0; i < n / 4; i += stride) p4[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    for (long long i = (n / 4) * 4 + i0; i < n; i += stride) p[i] = 0.f;
}

// The step's prologue, grid-stride: sigmoid of the opacity logits (when
// logits is given; the accumulated opacities come from k_accum_sigmoid
// instead), zeroed gradient accumulators and per-triangle counts, clip =
// M_b [x, y, z, 1] for every (view, vertex) (row-major M,
// examples/utils.py:412-426, summed in the order ((m0 x + m1 y) + m2 z) + m3),
// the raster bins' cleared tile counters and stats block, and the shading
// kernel's weight operands.
__global__ void k_step_prep(const float* __restrict__ logits, float* __restrict__ alpha, long long n_alpha,
                            float* __restrict__ galpha, float* __restrict__ gpos, long long n_pos,
                            int* __restrict__ tri_count, int T, float* __restrict__ gfeat, long long n_feat,
                            float* __restrict__ z1, long long n_z1, float* __restrict__ z2, long long n_z2,
                            int B, int V, const float* __restrict__ verts, const float* __restrict__ mvp,
                            float4* __restrict__ pos, int* __restrict__ tile_count, int n_tiles,
                            long long* __restrict__ stats, BlobJob blob) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    const long long i0 = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    for (long long i = i0; i < n_alpha; i += stride) {
        // evaluated in double and rounded once, so the opacity is the
        // correctly rounded sigmoid (the last bit decides mask ties)
        if (logits) alpha[i] = static_cast<float>(1.0 / (1.0 + exp(-static_cast<double>(logits[i]))));
        galpha[i] = 0.f;
    }
    zero_floats(gfeat, n_feat, i0, stride);
    zero_floats(z1, n_z1, i0, stride);
    zero_floats(z2, n_z2, i0, stride);
    float4* gp4 = reinterpret_cast<float4*>(gpos);
    for (long long i = i0; i < n_pos / 4; i += stride) gp4[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    for (long long i = i0; i < T; i += stride) tri_count[i] = 0;
    for (long long i = i0; i < (long long)B * V; i += stride) {
        const int b = static_cast<int>(i / V);
        const int v = static_cast<int>(i - (long long)b * V);
        const float x = verts[3 * v], y = verts[3 * v + 1], z = verts[3 * v + 2];
        const float* m = mvp + 16 * b;
        float o[4];
        for (int r = 0; r < 4; ++r) o[r] = m[4 * r] * x + m[4 * r + 1] * y + m[4 * r + 2] * z + m[4 * r + 3];
        pos[i] = make_float4(o[0], o[1], o[2], o[3]);
    }
    for (long long i = i0; i < n_tiles; i += stride) tile_count[i] = 0;
    if (i0 < kStatCount) stats[i0] = 0;
    for (long long i = i0; i < blob.n(); i += stride) build_blob(blob, static_cast<int>(i));
}

// One accumulation of the reference's composition: the level-summed lattice
// function of src [T, S_tot(lo..hi), C] evaluated at the vertices of the
// target lattice (accumulate_to_level, python/diffsoup/multires.py:279-339,
// src/cuda/multires.cu:314-389), then its sigmoid (examples/02_synthetic.py:
// 355-356): dst [T, S_T, C].  The backward reads dst (the sigmoid) and the
// gradient of dst and writes scale * (transpose of the accumulation)(g s (1 - s))
// into grad_src.
struct AccumJob {
    const float* src;
    float* dst;
    const float* grad_dst;   // backward
    float* grad_src;         // backward (every row written)
    float scale;             // backward
    int lo, hi, target, C, S_tot, S_T;
};

// Two source rows into LDS as one, n0 floats of s0 then n1 of s1 at dst,
// by loads that write LDS directly (global_load_lds_dword: lane l of a wave
// writes the dword at the wave's LDS base + 4 l, the base taken from the
// first active lane, so every lane passes its wave's base).  No registers
// hold the data, so all of a row's loads are in flight at once; the
// vector-memory counter is drained before the caller's barrier.
// blockDim.x must be a multiple of the wave.
__device__ __forceinline__ void stage_rows_lds(float* dst, const float* __restrict__ s0, int n0,
                                               const float* __restrict__ s1, int n1) {
    const int lane = threadIdx.x % kWave;
    for (int e = threadIdx.x; e < n0 + n1; e += blockDim.x) {
        const float* g = e < n0 ? s0 + e : s1 + (e - n0);
        __builtin_amdgcn_global_load_lds(g, dst + (e - lane), 4, 0, 0);
    }
    __builtin_amdgcn_s_waitcnt(0x0f70);   // vmcnt(0); expcnt, lgkmcnt not waited on
}

// The corner form of one accumulation, used for the opacities when their
// levels differ from the features': with the triangle's source row staged
// in LDS, each thread takes target samples, forms the sample's three corners
// at every level once and sums all C channels in the order of
// ds_accumulate_fwd.  Opacities go through a sigmoid evaluated in double and
// rounded once (their last bit decides mask ties, as in the step's
// prologue); features through the fp32 sigmoid (the reference applies
// torch's fp32 sigmoid).  C is a template argument so the index arithmetic
// folds; C = 0 is the runtime-width fallback.
template <int C>
__device__ __forceinline__ void accum_row_fwd(const AccumJob& J, int t, const float* row, bool exact_sigmoid) {
    const int c = C > 0 ? C : J.C;
    const int n_dst = J.S_T * c;
    const float Rf = static_cast<float>(1u << J.target);
    float* dst = J.dst + (long long)t * n_dst;
    constexpr int CA = C > 0 ? C : 16;
    for (int k = threadIdx.x; k < J.S_T; k += blockDim.x) {
        uint32_t x, y;
        lattice_xy(static_cast<uint32_t>(k), x, y);
        const float b0 = static_cast<float>(x) / Rf, b1 = static_cast<float>(y) / Rf;
        for (int j0 = 0; j0 < c; j0 += CA) {
            const int nc = min(CA, c - j0);
            float v[CA];
#pragma unroll
            for (int jj = 0; jj < CA; ++jj) v[jj] = 0.f;
            uint32_t off = 0;
            for (int l = J.lo; l <= J.hi; ++l) {
                // one cell per level, all channels summed in the order of the
                // single-channel loop: corner 0, 1, 2 of each level in turn
                const Corners cr = lattice_corners(b0, b1, static_cast<uint32_t>(l));
                const float* f0 = row + (off + cr.idx[0]) * c + j0;
                const float* f1 = row + (off + cr.idx[1]) * c + j0;
                const float* f2 = row + (off + cr.idx[2]) * c + j0;
#pragma unroll
                for (int jj = 0; jj < CA; ++jj)
                    if (C > 0 || jj < nc) {
                        v[jj] += cr.w[0] * f0[jj];
                        v[jj] += cr.w[1] * f1[jj];
                        v[jj] += cr.w[2] * f2[jj];
                    }
                off += feats_at_level(static_cast<uint32_t>(l));
            }
#pragma unroll
            for (int jj = 0; jj < CA; ++jj)
                if (C > 0 || jj < nc)
                    dst[k * c + j0 + jj] = exact_sigmoid
                                               ? static_cast<float>(1.0 / (1.0 + exp(-static_cast<double>(v[jj]))))
                                               : 1.0f / (1.0f + expf(-v[jj]));
        }
    }
}

// Forward of the feature accumulation by prolongation, in place in the
// staged source row: with r_lo = src_lo, r_(l+1) = src_(l+1) + P r_l, where P
// interpolates level l linearly onto level l + 1 (a level-(l+1) vertex is a
// level-l vertex, or the midpoint of a level-l edge: the average of its two
// ends, the cells split along the anti-diagonal as in lattice_corners).
// r_hi is the accumulation at the target vertices (the same hat-function sums
// as the corner form, grouped level by level), left contiguous in LDS, so
// the sigmoid's output row is stored in one coalesced pass.  A thread takes
// a fine vertex and all its channels; per level two LDS reads per channel at
// most, against three corners per level at every target vertex in the
// corner form.  With OP the opacity row (one channel, the same levels) goes
// through the same passes as one more channel, reusing each fine vertex's
// indices, and its sigmoid is evaluated in double and rounded once; without
// it the opacities take the corner form above.
template <int C, bool OP>
__device__ __forceinline__ void accum_feat_fwd(const AccumJob& J, int t, float* row, const AccumJob& A, float* arow) {
    const int c = C > 0 ? C : J.C;
    const int n_dst = J.S_T * c;
    constexpr int CA = C > 0 ? C : 16;
    int off = 0;   // first sample of level l in the row
    for (int l = J.lo; l < J.hi; ++l) {
        const int off_f = off + static_cast<int>(feats_at_level(static_cast<uint32_t>(l)));
        const int S_f = static_cast<int>(feats_at_level(static_cast<uint32_t>(l + 1)));
        for (int k = threadIdx.x; k < S_f; k += blockDim.x) {
            uint32_t X, Y;
            lattice_xy(static_cast<uint32_t>(k), X, Y);
            // the level-l vertices a (and b) whose average is P r_l here
            uint32_t ax, ay, bx, by;
            if (!(X & 1u) && !(Y & 1u)) { ax = bx = X >> 1; ay = by = Y >> 1; }
            else if (!(Y & 1u)) { ax = X >> 1; bx = ax + 1u; ay = by = Y >> 1; }       // X odd
            else if (!(X & 1u)) { ax = bx = X >> 1; ay = Y >> 1; by = ay + 1u; }       // Y odd
            else { ax = (X + 1u) >> 1; ay = Y >> 1; bx = X >> 1; by = (Y + 1u) >> 1; }  // both odd
            const uint32_t ka = (ax + ay) * (ax + ay + 1u) / 2u + ay, kb = (bx + by) * (bx + by + 1u) / 2u + by;
            const float* pa = row + (off + static_cast<int>(ka)) * c;
            const float* pb = row + (off + static_cast<int>(kb)) * c;
            float* q = row + (off_f + k) * c;
            for (int j0 = 0; j0 < c; j0 += CA) {
#pragma unroll
                for (int jj = 0; jj < CA; ++jj)
                    if (C > 0 || j0 + jj < c) q[j0 + jj] += 0.5f * (pa[j0 + jj] + pb[j0 + jj]);
            }
            if constexpr (OP) arow[off_f + k] += 0.5f * (arow[off + static_cast<int>(ka)] + arow[off + static_cast<int>(kb)]);
        }
        __syncthreads();
        off = off_f;
    }
    float* dst = J.dst + (long long)t * n_dst;
    const float* r = row + off * c;
    for (int i = threadIdx.x; i < n_dst; i += blockDim.x) dst[i] = 1.0f / (1.0f + expf(-r[i]));
    if constexpr (OP) {
        float* da = A.dst + (long long)t * A.S_T;
        for (int k = threadIdx.x; k < A.S_T; k += blockDim.x)
            da[k] = static_cast<float>(1.0 / (1.0 + exp(-static_cast<double>(arow[off + k]))));
    }
}

// Both source rows of a triangle are staged together (features, then the
// opacities after them), so a triangle waits out one round of loads.  With
// merge (the host's accum_merge: the opacities over the features' levels, 7
// or 8 feature channels) the opacities ride along as one more channel.
__global__ void __launch_bounds__(256) k_accum_sigmoid(int T, AccumJob f, AccumJob a, int merge) {
    extern __shared__ float lds_row[];
    const int nf = f.S_tot * f.C, na = a.S_tot;
    float* frow = lds_row;
    float* arow = lds_row + nf;
    const bool same = merge != 0;
    for (int t = blockIdx.x; t < T; t += gridDim.x) {
        stage_rows_lds(lds_row, f.src + (long long)t * nf, nf, a.src + (long long)t * na, na);
        __syncthreads();
        if (same) {
            if (f.C == 7) accum_feat_fwd<7, true>(f, t, frow, a, arow);
            else accum_feat_fwd<8, true>(f, t, frow, a, arow);
        } else {
            if (f.C == 7) accum_feat_fwd<7, false>(f, t, frow, a, arow);
            else if (f.C == 8) accum_feat_fwd<8, false>(f, t, frow, a, arow);
            else accum_feat_fwd<0, false>(f, t, frow, a, arow);
            accum_row_fwd<1>(a, t, arow, true);
        }
        __syncthreads();   // the next triangle's staging overwrites both rows
    }
}

// Backward, one workgroup per triangle, by restriction.  The weight of
// source vertex V of level l at a target vertex p is V's hat function on the
// level-l lattice (it equals the weight lattice_corners gives V wherever p
// falls), and the lattices are nested: level l + 1 splits every level-l cell
// at its edge midpoints, so phi^l_V = phi^(l+1)_2V + 1/2 sum of phi^(l+1)_N
// over the six level-(l+1) neighbours N of 2V (the midpoints of V's edges).
// Hence, with r_hi = g = scale grad_dst s (1 - s) at the target vertices,
// grad_src at level l is r_l[V] = r_(l+1)[2V] + 1/2 sum_N r_(l+1)[N], one
// level at a time from the top: seven LDS reads per coarse vertex and
// channel instead of a walk over every target vertex in V's support, no
// atomics, and every source row written once.  (The sums are regrouped with
// respect to the scatter of ds_accumulate_bwd; the arithmetic is otherwise
// the same, the halves exact.)
template <int C>
__device__ __forceinline__ void accum_row_bwd(const AccumJob& J, int t, float* lds) {
    const int c = C > 0 ? C : J.C;
    const int n_dst = J.S_T * c;
    const float* gd = J.grad_dst + (long long)t * n_dst;
    const float* sd = J.dst + (long long)t * n_dst;
    float* fine = lds;                 // r of the finer level, [S_(l+1), c]
    float* coarse = lds + n_dst;       // r of level l, [S_l, c]
    {
        constexpr int U = 4;   // loads in flight per thread: 2 U
        const int nt = blockDim.x;
        int i = threadIdx.x;
        for (; i + (U - 1) * nt < n_dst; i += U * nt) {
            float a[U], b[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                a[u] = gd[i + u * nt];
                b[u] = sd[i + u * nt];
            }
#pragma unroll
            for (int u = 0; u < U; ++u) fine[i + u * nt] = J.scale * a[u] * (b[u] * (1.0f - b[u]));
        }
        for (; i < n_dst; i += nt) {
            const float sg = sd[i];
            fine[i] = J.scale * gd[i] * (sg * (1.0f - sg));
        }
    }
    __syncthreads();
    // offsets of the source levels in a row
    int off_hi = 0;
    for (int l = J.lo; l < J.hi; ++l) off_hi += static_cast<int>(feats_at_level(static_cast<uint32_t>(l)));
    float* out = J.grad_src + (long long)t * J.S_tot * c;
    // level hi is the identity (the target lattice is level hi)
    for (int i = threadIdx.x; i < n_dst; i += blockDim.x) out[off_hi * c + i] = fine[i];
    int off = off_hi;
    for (int l = J.hi - 1; l >= J.lo; --l) {
        const int S = static_cast<int>(feats_at_level(static_cast<uint32_t>(l)));
        const int Rf = 1 << (l + 1);   // the finer level's edge
        off -= S;
        for (int u = threadIdx.x; u < S * c; u += blockDim.x) {
            const int k = u / c, j = u - k * c;
            uint32_t X, Y;
            lattice_xy(static_cast<uint32_t>(k), X, Y);
            const int x = 2 * static_cast<int>(X), y = 2 * static_cast<int>(Y);
            auto at = [&](int px, int py) -> float {
                if (px < 0 || py < 0 || px + py > Rf) return 0.f;
                const int n = px + py;
                return fine[(((n * (n + 1)) >> 1) + py) * c + j];
            };
            const float nb = ((at(x + 1, y) + at(x - 1, y)) + (at(x, y + 1) + at(x, y - 1))) +
                             (at(x + 1, y - 1) + at(x - 1, y + 1));
            coarse[u] = at(x, y) + 0.5f * nb;
        }
        __syncthreads();
        for (int u = threadIdx.x; u < S * c; u += blockDim.x) {
            out[off * c + u] = coarse[u];
            fine[u] = coarse[u];
        }
        __syncthreads();
    }
    // with a single level only the identity copy read the row: order those
    // reads before the next job's (or triangle's) writes of it
    if (J.lo == J.hi) __syncthreads();
}

// The same restriction for the features (C channels) and the opacities (one
// channel over the same levels) together, vertex-major: in LDS a vertex holds
// CP = C + 1 values rounded up to whole float4s, the opacity after the
// features, and a thread takes one coarse vertex at a time, forms its seven
// fine indices once and reads each fine vertex as float4s for every channel.
// Each value is formed by the arithmetic of accum_row_bwd in the same order.
// The top level's gradient is stored from registers as it is staged, each
// coarse level as it is formed, and the levels alternate between two LDS
// regions, so a level costs one barrier.
template <int C>
__device__ __forceinline__ void accum_rows_bwd_op(const AccumJob& F, const AccumJob& A, int t, float* lds) {
    constexpr int CP = (C + 1 + 3) & ~3, NQ = CP / 4;
    const int S_T = F.S_T, n_f = S_T * C, n = n_f + S_T;
    int off_hi = 0;   // first sample of level hi in a row
    for (int l = F.lo; l < F.hi; ++l) off_hi += static_cast<int>(feats_at_level(static_cast<uint32_t>(l)));
    const float* gdf = F.grad_dst + (long long)t * n_f;
    const float* sdf = F.dst + (long long)t * n_f;
    const float* gda = A.grad_dst + (long long)t * S_T;
    const float* sda = A.dst + (long long)t * S_T;
    float* outf = F.grad_src + (long long)t * F.S_tot * C;
    float* outa = A.grad_src + (long long)t * A.S_tot;
    float* fine = lds;                   // [S_T][CP]
    float* coarse = lds + S_T * CP;      // [S_(hi-1)][CP]
    {
        constexpr int U = 8;   // loads in flight per thread: 2 U
        const int nt = blockDim.x;
        auto put = [&](int e, float ge, float se) {
            if (e < n_f) {
                const float r = F.scale * ge * (se * (1.0f - se));
                fine[(e / C) * CP + e % C] = r;
                outf[off_hi * C + e] = r;
            } else {
                const float r = A.scale * ge * (se * (1.0f - se));
                fine[(e - n_f) * CP + C] = r;
                outa[off_hi + e - n_f] = r;
            }
        };
        int i = threadIdx.x;
        for (; i + (U - 1) * nt < n; i += U * nt) {
            float a[U], b[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int e = i + u * nt;
                a[u] = e < n_f ? gdf[e] : gda[e - n_f];
                b[u] = e < n_f ? sdf[e] : sda[e - n_f];
            }
#pragma unroll
            for (int u = 0; u < U; ++u) put(i + u * nt, a[u], b[u]);
        }
        for (; i < n; i += nt) put(i, i < n_f ? gdf[i] : gda[i - n_f], i < n_f ? sdf[i] : sda[i - n_f]);
    }
    __syncthreads();
    int off = off_hi;
    for (int l = F.hi - 1; l >= F.lo; --l) {
        const int S = static_cast<int>(feats_at_level(static_cast<uint32_t>(l)));
        const int Rf = 1 << (l + 1);   // the finer level's edge
        off -= S;
        for (int k = threadIdx.x; k < S; k += blockDim.x) {
            uint32_t X, Y;
            lattice_xy(static_cast<uint32_t>(k), X, Y);
            const int x = 2 * static_cast<int>(X), y = 2 * static_cast<int>(Y);
            // fine vertex index of (px, py), or -1 outside the lattice
            auto idx = [&](int px, int py) -> int {
                if (px < 0 || py < 0 || px + py > Rf) return -1;
                const int m = px + py;
                return ((m * (m + 1)) >> 1) + py;
            };
            const int i0 = idx(x, y), i1 = idx(x + 1, y), i2 = idx(x - 1, y), i3 = idx(x, y + 1),
                      i4 = idx(x, y - 1), i5 = idx(x + 1, y - 1), i6 = idx(x - 1, y + 1);
            auto row = [&](int v, int q) -> float4 {
                return v < 0 ? make_float4(0.f, 0.f, 0.f, 0.f) : *reinterpret_cast<const float4*>(fine + v * CP + 4 * q);
            };
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const float4 c0 = row(i0, q), c1 = row(i1, q), c2 = row(i2, q), c3 = row(i3, q), c4 = row(i4, q),
                             c5 = row(i5, q), c6 = row(i6, q);
                float r[4];
                const float v0[4] = {c0.x, c0.y, c0.z, c0.w}, v1[4] = {c1.x, c1.y, c1.z, c1.w},
                            v2[4] = {c2.x, c2.y, c2.z, c2.w}, v3[4] = {c3.x, c3.y, c3.z, c3.w},
                            v4[4] = {c4.x, c4.y, c4.z, c4.w}, v5[4] = {c5.x, c5.y, c5.z, c5.w},
                            v6[4] = {c6.x, c6.y, c6.z, c6.w};
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const float nb = ((v1[e] + v2[e]) + (v3[e] + v4[e])) + (v5[e] + v6[e]);
                    r[e] = v0[e] + 0.5f * nb;
                }
                *reinterpret_cast<float4*>(coarse + k * CP + 4 * q) = make_float4(r[0], r[1], r[2], r[3]);
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int j = 4 * q + e;
                    if (j < C) outf[(off + k) * C + j] = r[e];
                    else if (j == C) outa[off + k] = r[e];
                }
            }
        }
        __syncthreads();
        float* tmp = fine;
        fine = coarse;
        coarse = tmp;
    }
}

__global__ void __launch_bounds__(256) k_accum_sigmoid_bwd(int T, AccumJob f, AccumJob a, int merge) {
    extern __shared__ float lds_row[];
    const bool same = merge != 0;
    for (int t = blockIdx.x; t < T; t += gridDim.x) {
        if (same) {
            if (f.C == 7) accum_rows_bwd_op<7>(f, a, t, lds_row);
            else accum_rows_bwd_op<8>(f, a, t, lds_row);
            continue;
        }
        if (f.C == 7) accum_row_bwd<7>(f, t, lds_row);
        else if (f.C == 8) accum_row_bwd<8>(f, t, lds_row);
        else accum_row_bwd<0>(f, t, lds_row);
        accum_row_bwd<1>(a, t, lds_row);
    }
}

// The step's closing reductions in one launch of 1024-thread blocks:
//   blocks [0, R)       grad_mlp[k] = sum over shading blocks of partials[.][k],
//                       64 parameters per block, 16 partial rows per pass;
//   block R             loss = (sum of the per-tile background losses + the
//                       shaded-pixel loss column of the partials) * scale;
//   block R + 1         stats[covered] = sum of the per-tile covered counts;
//   blocks (R + 1, R + 1 + F]
//                       grad_verts[v] = sum_b (M_b^T grad_pos[b, v])[0:3] and
//                       grad_logits = scale * g * s (1 - s), s = sigmoid.
__global__ void __launch_bounds__(1024) k_epilogue(const float* __restrict__ partials, int nblocks, int nparam,
                                                   float* __restrict__ grad_mlp, const float* __restrict__ bg,
                                                   int n_bg, float scale, float* __restrict__ loss, int B, int V,
                                                   const float* __restrict__ mvp, const float4* __restrict__ gpos,
                                                   float* __restrict__ gverts, long long n_alpha,
                                                   const float* __restrict__ alpha,
                                                   const float* __restrict__ galpha, float* __restrict__ glogits,
                                                   const int* __restrict__ tile_cov, int n_tiles,
                                                   long long* __restrict__ stats,
                                                   unsigned long long* __restrict__ seed_advance,
                                                   long long* __restrict__ counter_advance) {
    __shared__ float acc[16][65];
    __shared__ long long cov[16];
    const int R = (nparam + 63) / 64;
    const int tx = threadIdx.x % 64, ty = threadIdx.x / 64;
    if ((int)blockIdx.x < R) {
        const int k = blockIdx.x * 64 + tx;
        // four independent sums, so a thread has four loads in flight
        float s4[4] = {0.f, 0.f, 0.f, 0.f};
        if (k < nparam) {
            int b = ty;
            for (; b + 48 < nblocks; b += 64)
#pragma unroll
                for (int u = 0; u < 4; ++u) s4[u] += partials[(long long)(b + 16 * u) * nparam + k];
            for (; b < nblocks; b += 16) s4[0] += partials[(long long)b * nparam + k];
        }
        const float sum = (s4[0] + s4[1]) + (s4[2] + s4[3]);
        acc[ty][tx] = sum;
        __syncthreads();
        if (ty == 0 && k < nparam) {
            float t = 0.f;
            for (int r = 0; r < 16; ++r) t += acc[r][tx];
            grad_mlp[k] = t;
        }
        return;
    }
    if ((int)blockIdx.x == R) {
        float s4[4] = {0.f, 0.f, 0.f, 0.f};
        {
            int i = threadIdx.x;
            for (; i + 3 * (int)blockDim.x < n_bg; i += 4 * blockDim.x)
#pragma unroll
                for (int u = 0; u < 4; ++u) s4[u] += bg[i + u * blockDim.x];
            for (; i < n_bg; i += blockDim.x) s4[0] += bg[i];
        }
        float sum = (s4[0] + s4[1]) + (s4[2] + s4[3]);
        for (int b = threadIdx.x; b < nblocks; b += blockDim.x) sum += partials[(long long)b * nparam + nparam - 1];
        for (int d = 32; d > 0; d >>= 1) sum += __shfl_xor(sum, d, 64);
        if (tx == 0) acc[0][ty] = sum;
        __syncthreads();
        if (threadIdx.x == 0 && loss) {   // null when another kernel forms the loss
            float t = 0.f;
            for (int q = 0; q < 16; ++q) t += acc[0][q];
            *loss = t * scale;
        }
        return;
    }
    if ((int)blockIdx.x == R + 1) {
        long long c4[4] = {0, 0, 0, 0};
        {
            int i = threadIdx.x;
            for (; i + 3 * (int)blockDim.x < n_tiles; i += 4 * blockDim.x)
#pragma unroll
                for (int u = 0; u < 4; ++u) c4[u] += tile_cov[i + u * blockDim.x];
            for (; i < n_tiles; i += blockDim.x) c4[0] += tile_cov[i];
        }
        long long c = (c4[0] + c4[1]) + (c4[2] + c4[3]);
        for (int d = 32; d > 0; d >>= 1) c += __shfl_xor(c, d, 64);
        if (tx == 0) cov[ty] = c;
        __syncthreads();
        if (threadIdx.x == 0) {
            long long t = 0;
            for (int q = 0; q < 16; ++q) t += cov[q];
            stats[kStatCovered] = t;
            // the mask seed's last reader (the resolve) has run
            if (seed_advance) *seed_advance += 1ull;
            // the optimiser that follows reads its counter as advanced
            if (counter_advance) *counter_advance += 1ll;
        }
        return;
    }
    const long long id = ((long long)blockIdx.x - R - 2) * blockDim.x + threadIdx.x;
    if (id < V) {
        float gx = 0.f, gy = 0.f, gz = 0.f;
        for (int b = 0; b < B; ++b) {
            const float4 g = gpos[(long long)b * V + id];
            const float* m = mvp + 16 * b;
            gx += m[0] * g.x + m[4] * g.y + m[8] * g.z + m[12] * g.w;
            gy += m[1] * g.x + m[5] * g.y + m[9] * g.z + m[13] * g.w;
            gz += m[2] * g.x + m[6] * g.y + m[10] * g.z + m[14] * g.w;
        }
        gverts[3 * id] = gx;
        gverts[3 * id + 1] = gy;
        gverts[3 * id + 2] = gz;
    }
    if (id < n_alpha) {
        const float sa = alpha[id];
        glogits[id] = scale * galpha[id] * (sa * (1.0f - sa));
    }
}

inline unsigned blocks_for(long long n, int bs) { return static_cast<unsigned>((n + bs - 1) / bs); }
inline size_t up256(size_t x) { return (x + 255) & ~size_t(255); }

struct StepLayout {
    size_t raster, gpos, galpha, tile_list, tile_ncov, tri_count, tri_off, list, scan, bgloss, blobs, partials,
        tile_px_off, px_list, scan2, photo, feat_acc, gfeat_acc, bounds, aux_rec, aux_n,
        total;
    int n_tiles, nblocks, nparam;
    int nl;          // levels the shading kernel gathers (1 with accumulate)
    bool sh;         // SH2 inputs
    bool direct;     // the shading reads and scatters its rows in global memory
    long long P;
    long long S_src;   // samples per triangle of d->feat
    long long S_shade; // samples per triangle of the table the shading reads
    long long Sa_src;  // samples per triangle of d->alpha_logits
    long long Sa;      // opacity samples at alpha_level
    int aux_per_quad;  // fragment records per 8x8 quad for the opacity gradient
};

// Compute units of the current device (256 on MI355X when the query fails).
int device_cus() {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n < 1)
        return 256;
    return n;
}

// Shading workgroups: one round of resident workgroups (per_cu on every
// compute unit), each wave taking one cost-balanced range; fewer for small
// batches (64 covered pixels per workgroup at least).
int shade_blocks(long long P, int per_cu) {
    long long nb = (P + 63) / 64;
    const long long cap = static_cast<long long>(device_cus()) * per_cu;
    if (nb > cap) nb = cap;
    if (nb < 1) nb = 1;
    return static_cast<int>(nb);
}

bool layout(const ds_step_desc* d, const Views& v, StepLayout& L) {
    const bool acc = d->accumulate != 0;
    L.nl = acc ? 1 : d->feat_hi - d->feat_lo + 1;
    L.sh = d->sh2 != 0;
    if (!shade_supported(L.nl, d->channels, d->hidden, L.sh)) return false;
    L.n_tiles = v.tile_base[v.n];
    L.P = v.pix_base[v.n];
    L.nblocks = shade_blocks(L.P, shade_blocks_per_cu(L.nl, d->channels, d->hidden, L.sh));
    L.nparam = shade_nparam(L.nl, d->channels, d->hidden, L.sh);
    L.Sa = feats_at_level(d->alpha_level);
    L.Sa_src = acc ? feats_in_levels(d->alpha_lo, d->alpha_level) : L.Sa;
    L.S_src = feats_in_levels(d->feat_lo, d->feat_hi);
    L.S_shade = acc ? feats_at_level(d->feat_hi) : L.S_src;
    size_t o = 0;
    L.raster = o; o += up256(raster_workspace_bytes(d->B, d->T, L.n_tiles, d->capacity));
    L.gpos = o; o += up256(sizeof(float) * 4 * (size_t)d->B * d->V);
    L.galpha = o; o += up256(sizeof(float) * (size_t)d->T * L.Sa);
    const size_t nt2 = (size_t)d->tile * d->tile;
    L.tile_list = o; o += up256(sizeof(int) * 2 * (size_t)L.n_tiles * nt2);
    L.tile_ncov = o; o += up256(sizeof(int) * (size_t)L.n_tiles);
    L.tri_count = o; o += up256(sizeof(int) * (size_t)(d->T > 0 ? d->T : 1));
    L.tri_off = o; o += up256(sizeof(int) * ((size_t)d->T + 1));
    L.list = o; o += up256(sizeof(PixRec) * (size_t)L.P);
    L.scan = o; o += up256(scan_tmp_bytes(d->T));
    L.tile_px_off = o; o += up256(sizeof(int) * ((size_t)L.n_tiles + 1));
    L.px_list = o; o += up256(sizeof(int4) * (size_t)L.P);
    L.scan2 = o; o += up256(scan_tmp_bytes(L.n_tiles));
    L.bgloss = o; o += up256(sizeof(float) * (size_t)L.n_tiles);
    L.blobs = o; o += up256(sizeof(float) * shade_blob_floats(L.nl, d->channels, d->hidden, L.sh));
    L.partials = o; o += up256(sizeof(float) * (size_t)L.nblocks * L.nparam);
    L.bounds = o; o += up256(sizeof(int) * ((size_t)L.nblocks * 4 + 1));
    L.aux_per_quad = d->aux_records > 0 ? static_cast<int>(std::min<int64_t>(d->aux_records, 1 << 16)) : 128;
    const size_t n_quads = (size_t)L.n_tiles * nt2 / 64;
    L.aux_rec = o; o += up256(sizeof(int4) * n_quads * L.aux_per_quad);
    L.aux_n = o; o += up256(sizeof(int) * n_quads);
    L.direct = shade_direct_rows(L.nl, d->channels, d->hidden, L.sh, d->bf16 != 0,
                                 !acc && d->feat_bf16 != nullptr, static_cast<int>(L.S_shade));
    L.feat_acc = L.gfeat_acc = o;
    if (acc) {
        const size_t nacc = sizeof(float) * (size_t)d->T * L.S_shade * d->channels;
        L.feat_acc = o; o += up256(nacc);
        L.gfeat_acc = o; o += up256(nacc);
    }
    L.photo = o;
    if (d->loss_kind == DS_LOSS_L1_SSIM) {
        const size_t nf = photo_loss_floats(v, 3);
        if (nf == 0) return false;
        o += up256(sizeof(float) * nf);
    }
    L.total = o;
    return true;
}

// Everything one enqueue of the step needs: the views, the workspace layout
// and its carve-up.
struct Step {
    Views v;
    StepLayout L;
    RasterWorkspace rw;
    float *gpos, *galpha, *bgloss, *blobs, *partials, *feat_acc, *gfeat_acc, *photo;
    int2* tile_list;
    int *tile_ncov, *tri_count, *tri_off, *tile_px_off, *bounds;
    PixRec* list;
    int4 *px_list, *aux_rec;
    int* aux_n;
    void *scan_tmp, *scan_tmp2;
    float inv_n;   // 1 / (3 P): the MSE mean and the opacity term's 1 / numel(color)
    hipStream_t s;
    // the shading launch took the visible fragments' opacity gradient: the
    // opacity-gradient pass then skips them
    bool va = false;
};

int step_prepare(const ds_step_desc* d, void* stream, Step& st);
size_t accum_lds_bytes(const ds_step_desc* d, const StepLayout& L);
int enqueue_forward(const ds_step_desc* d, Step& st, bool clear_rows);
int enqueue_shade(const ds_step_desc* d, Step& st, int fwd_only, const float* ext_grad);
int enqueue_backward(const ds_step_desc* d, Step& st, const float* grad_color, bool write_loss);

}  // namespace

// Exported through capi.hip's error helpers.
int step_make_views(const ds_step_desc* d, Views& v);
int step_fail(int code, const char* msg);

namespace {

int step_prepare(const ds_step_desc* d, void* stream, Step& st) {
    if (!d) return step_fail(DS_ERR_INVALID, "null step descriptor");
    int rc = step_make_views(d, st.v);
    if (rc != DS_OK) return rc;
    const int nl = d->feat_hi - d->feat_lo + 1;
    if (d->feat_lo < 0 || nl < 1 || d->feat_hi > 10 || d->alpha_level < 0 || d->alpha_level > 10)
        return step_fail(DS_ERR_INVALID, "level range invalid");
    if (d->accumulate && (d->alpha_lo < 0 || d->alpha_lo > d->alpha_level))
        return step_fail(DS_ERR_INVALID, "alpha_lo must be in [0, alpha_level]");
    if (d->accumulate && d->feat_bf16)
        return step_fail(DS_ERR_INVALID, "accumulated features are fp32: feat_bf16 must be null");
    if (d->sh2 && !d->inv_mvp) return step_fail(DS_ERR_INVALID, "sh2 needs inv_mvp");
    if (d->T < 0 || d->V < 0 || d->T > (1 << 24) - 1) return step_fail(DS_ERR_INVALID, "bad T or V");
    if (d->loss_kind != DS_LOSS_MSE && d->loss_kind != DS_LOSS_L1_SSIM)
        return step_fail(DS_ERR_INVALID, "loss_kind must be DS_LOSS_MSE or DS_LOSS_L1_SSIM");
    StepLayout& L = st.L;
    if (!layout(d, st.v, L))
        return step_fail(DS_ERR_UNSUPPORTED, "no compiled shading kernel for this (levels, channels, hidden, sh2)");
    if (d->accumulate && accum_lds_bytes(d, L) > 160 * 1024)
        return step_fail(DS_ERR_UNSUPPORTED, "accumulate: a triangle's two source rows (levels x channels, and "
                                             "the opacity levels) must fit 160 KB of LDS (levels up to 6 with 7 "
                                             "channels do)");
    if (!d->workspace || d->workspace_bytes < L.total)
        return step_fail(DS_ERR_WORKSPACE, "workspace smaller than ds_step_workspace_size");
    if (reinterpret_cast<uintptr_t>(d->workspace) & 255u) return step_fail(DS_ERR_INVALID, "workspace must be 256-byte aligned");
    const void* feat_in = d->feat_bf16 ? static_cast<const void*>(d->feat_bf16) : static_cast<const void*>(d->feat);
    const void* req[] = {d->verts, d->mvp, d->tri, d->alpha_logits, feat_in, d->w1, d->b1, d->w2, d->b2,
                         d->w3, d->b3, d->target, d->pos, d->alpha, d->rast, d->color, d->grad_color,
                         d->pix_loss, d->grad_verts, d->grad_alpha_logits, d->grad_feat, d->grad_mlp, d->loss};
    for (const void* p : req)
        if (!p && d->T > 0) return step_fail(DS_ERR_INVALID, "null pointer in step descriptor");
    if (d->feat_bf16 && d->bf16 && d->T > 0 &&
        (!d->feat || (reinterpret_cast<uintptr_t>(d->feat) & 15u)))
        return step_fail(DS_ERR_INVALID, "bf16 feature rows also need the 16-byte aligned fp32 master feat "
                                         "(the residual-blend channels are read from it)");
    if ((reinterpret_cast<uintptr_t>(d->pos) | reinterpret_cast<uintptr_t>(d->rast) |
         reinterpret_cast<uintptr_t>(feat_in) | reinterpret_cast<uintptr_t>(d->grad_feat)) & 15u)
        return step_fail(DS_ERR_INVALID, "pos, rast, feat (or feat_bf16) and grad_feat must be 16-byte aligned");

    st.s = static_cast<hipStream_t>(stream);
    char* ws = static_cast<char*>(d->workspace);
    st.rw = raster_workspace_carve(ws + L.raster, d->B, d->T, L.n_tiles, d->capacity);
    st.gpos = reinterpret_cast<float*>(ws + L.gpos);
    st.galpha = reinterpret_cast<float*>(ws + L.galpha);
    st.tile_list = reinterpret_cast<int2*>(ws + L.tile_list);
    st.tile_ncov = reinterpret_cast<int*>(ws + L.tile_ncov);
    st.tri_count = reinterpret_cast<int*>(ws + L.tri_count);
    st.tri_off = reinterpret_cast<int*>(ws + L.tri_off);
    st.list = reinterpret_cast<PixRec*>(ws + L.list);
    st.scan_tmp = ws + L.scan;
    st.tile_px_off = reinterpret_cast<int*>(ws + L.tile_px_off);
    st.px_list = reinterpret_cast<int4*>(ws + L.px_list);
    st.aux_rec = reinterpret_cast<int4*>(ws + L.aux_rec);
    st.aux_n = reinterpret_cast<int*>(ws + L.aux_n);
    st.scan_tmp2 = ws + L.scan2;
    st.bgloss = reinterpret_cast<float*>(ws + L.bgloss);
    st.blobs = reinterpret_cast<float*>(ws + L.blobs);
    st.partials = reinterpret_cast<float*>(ws + L.partials);
    st.feat_acc = d->accumulate ? reinterpret_cast<float*>(ws + L.feat_acc) : nullptr;
    st.gfeat_acc = d->accumulate ? reinterpret_cast<float*>(ws + L.gfeat_acc) : nullptr;
    st.photo = reinterpret_cast<float*>(ws + L.photo);
    st.bounds = reinterpret_cast<int*>(ws + L.bounds);
    st.inv_n = 1.0f / (3.0f * static_cast<float>(L.P));
    return DS_OK;
}

#define DS_TRY(expr, what) \
    if ((e = (expr)) != hipSuccess) return step_fail(DS_ERR_HIP, what);

hipError_t mark(const ds_step_desc* d, hipStream_t s, int stage, bool begin) {
    if (d->timed_stage != stage) return hipSuccess;
    void* ev = begin ? d->ev_begin : d->ev_end;
    return ev ? hipEventRecord(static_cast<hipEvent_t>(ev), s) : hipSuccess;
}

// LDS of the accumulation kernels: both source rows (forward), or the
// target row plus the level below it of the larger job (backward).
size_t accum_lds_bytes(const ds_step_desc* d, const StepLayout& L, bool merge) {
    const size_t below = d->feat_hi > 0 ? feats_at_level(d->feat_hi - 1) : 0;
    const size_t below_a = d->alpha_level > 0 ? feats_at_level(d->alpha_level - 1) : 0;
    const size_t bwd = merge ? (size_t)(L.S_shade + below) * ((d->channels + 4) & ~3)
                             : std::max<size_t>((size_t)(L.S_shade + below) * d->channels, (size_t)L.Sa + below_a);
    return sizeof(float) * std::max<size_t>((size_t)L.S_src * d->channels + (size_t)L.Sa_src, bwd);
}

// Whether the accumulation kernels carry the opacities as one more channel
// of the features' passes: the same source levels and target level, a
// feature width those passes are compiled for (7 or 8), and their backward
// row within the LDS.
bool accum_merge(const ds_step_desc* d, const StepLayout& L) {
    return d->alpha_lo == d->feat_lo && d->alpha_level == d->feat_hi && L.Sa == L.S_shade && L.Sa_src == L.S_src &&
           (d->channels == 7 || d->channels == 8) && accum_lds_bytes(d, L, true) <= 160 * 1024;
}

size_t accum_lds_bytes(const ds_step_desc* d, const StepLayout& L) {
    return accum_lds_bytes(d, L, accum_merge(d, L));
}

// Rows beyond the default 64 KB of LDS need each kernel's opt-in, once.
hipError_t accum_lds_opt_in() {
    static bool done = false;
    if (done) return hipSuccess;
    for (const void* k : {reinterpret_cast<const void*>(k_accum_sigmoid),
                          reinterpret_cast<const void*>(k_accum_sigmoid_bwd)}) {
        const hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        if (e != hipSuccess) return e;
    }
    done = true;
    return hipSuccess;
}

AccumJob feat_job(const ds_step_desc* d, const Step& st) {
    return AccumJob{d->feat, st.feat_acc, st.gfeat_acc, d->grad_feat, 1.0f, d->feat_lo, d->feat_hi, d->feat_hi,
                    d->channels, static_cast<int>(st.L.S_src), static_cast<int>(st.L.S_shade)};
}

AccumJob alpha_job(const ds_step_desc* d, const Step& st) {
    return AccumJob{d->alpha_logits, d->alpha, st.galpha, d->grad_alpha_logits, st.inv_n, d->alpha_lo,
                    d->alpha_level, d->alpha_level, 1, static_cast<int>(st.L.Sa_src), static_cast<int>(st.L.Sa)};
}

// The shading ranges' job description: wave boundaries, and (clear_rows) the
// gradient rows the shading does not store whole.
ShadeRanges shade_ranges(const ds_step_desc* d, const Step& st, bool clear_rows) {
    const StepLayout& L = st.L;
    const bool acc = d->accumulate != 0;
    return ShadeRanges{st.tri_off, d->T, L.nblocks * kShadeWavesPerBlock, st.bounds,
                       clear_rows && !L.direct ? (acc ? st.gfeat_acc : d->grad_feat) : nullptr,
                       static_cast<int>(L.S_shade) * d->channels,
                       shade_seg_cost(L.nl, d->channels, d->hidden, L.sh)};
}

// Prologue, binning, resolve (with background shading) and the covered-pixel
// lists with the shading ranges: everything before the shading kernel.
int enqueue_forward(const ds_step_desc* d, Step& st, bool clear_rows) {
    const StepLayout& L = st.L;
    const hipStream_t s = st.s;
    const bool acc = d->accumulate != 0;
    const long long n_alpha = (long long)d->T * L.Sa;
    const long long n_feat = (long long)d->T * L.S_shade * d->channels;   // the table the shading writes
    const long long n_pos = 4LL * d->B * d->V;
    hipError_t e;
    const int din = L.nl * d->channels + (L.sh ? 9 : 0);
    const BlobJob blob{d->w1, d->w2, d->w3, st.blobs, (din + 15) / 16, d->hidden / 16, din, d->hidden, d->bf16 != 0};
    const long long prep_n = std::max(std::max(std::max(n_alpha, (long long)d->T), (long long)L.n_tiles),
                                      std::max(std::max(n_pos / 4, L.direct ? n_feat / 4 : 0LL),
                                               std::max(std::max((long long)d->B * d->V, (long long)kStatCount),
                                                        blob.n())));
    DS_TRY(mark(d, s, DS_STAGE_PREP, true), "event");
    k_step_prep<<<std::min<unsigned>(blocks_for(prep_n, 256), 4096), 256, 0, s>>>(
        acc ? nullptr : d->alpha_logits, d->alpha, n_alpha, st.galpha, st.gpos, n_pos, st.tri_count, d->T,
        acc ? st.gfeat_acc : d->grad_feat, L.direct ? n_feat : 0, nullptr, 0, nullptr, 0, d->B, d->V, d->verts,
        d->mvp,
        reinterpret_cast<float4*>(d->pos), st.rw.tile_count, L.n_tiles, st.rw.stats, blob);
    DS_TRY(hipGetLastError(), "step prep");
    if (acc && d->T > 0) {
        DS_TRY(accum_lds_opt_in(), "accumulate LDS");
        k_accum_sigmoid<<<std::min(d->T, 4096), 256, accum_lds_bytes(d, L), s>>>(
            d->T, feat_job(d, st), alpha_job(d, st), accum_merge(d, L) ? 1 : 0);
        DS_TRY(hipGetLastError(), "accumulate + sigmoid");
    }
    DS_TRY(mark(d, s, DS_STAGE_PREP, false), "event");
    DS_TRY(mark(d, s, DS_STAGE_BLOBS, true), "event");   // the blobs are built in the prologue
    DS_TRY(mark(d, s, DS_STAGE_BLOBS, false), "event");
    DS_TRY(mark(d, s, DS_STAGE_PROJECT, true), "event");   // projection runs in the prologue
    DS_TRY(mark(d, s, DS_STAGE_PROJECT, false), "event");
    DS_TRY(mark(d, s, DS_STAGE_BIN, true), "event");
    DS_TRY(launch_raster_bin(st.v, d->V, d->pos, d->T, d->tri, st.rw, s, /*clear=*/false), "raster bin");
    DS_TRY(mark(d, s, DS_STAGE_BIN, false), "event");
    AlphaSource a{d->alpha, d->alpha_level, static_cast<int>(L.Sa)};
    BackgroundShade bg{d->target, d->color, d->grad_color, d->pix_loss, st.bgloss, st.tile_list, st.tile_ncov,
                       st.tri_count, 2.0f * st.inv_n, d->background, st.aux_rec, st.aux_n, L.aux_per_quad};
    DS_TRY(mark(d, s, DS_STAGE_RESOLVE, true), "event");
    const MaskSeed seed{d->seed, reinterpret_cast<const unsigned long long*>(d->seed_device)};
    DS_TRY(launch_raster_resolve_shade(st.v, d->T, a, d->stochastic ? kMaskHash : kMaskDeterministic, seed,
                                       d->rast, st.rw, bg, s),
           "resolve");
    DS_TRY(mark(d, s, DS_STAGE_RESOLVE, false), "event");
    DS_TRY(launch_exclusive_scan_pair(st.tri_count, d->T, st.tri_off, st.scan_tmp, st.tile_ncov, L.n_tiles,
                                      st.tile_px_off, st.scan_tmp2, s),
           "triangle / tile scans");
    DS_TRY(launch_fill_triangle_lists(st.v, st.tile_list, st.tile_ncov, d->rast, d->target, st.tri_off, st.list,
                                      st.tile_px_off, st.px_list, shade_ranges(d, st, clear_rows), s),
           "triangle lists and shading ranges");
    return DS_OK;
}

// One shading launch: fused (colour, MSE gradient, adjoint), colour only, or
// the adjoint on an external colour gradient.
int enqueue_shade(const ds_step_desc* d, Step& st, int fwd_only, const float* ext_grad) {
    const StepLayout& L = st.L;
    const bool acc = d->accumulate != 0;
    ShadeArgs sa{reinterpret_cast<const float4*>(d->rast), st.tri_off, st.list, d->T, acc ? st.feat_acc : d->feat,
                 acc ? nullptr : reinterpret_cast<const unsigned short*>(d->feat_bf16),
                 acc ? st.gfeat_acc : d->grad_feat, static_cast<int>(L.S_shade), acc ? d->feat_hi : d->feat_lo,
                 st.blobs, d->b1, d->b2, d->b3, d->color, d->grad_color, d->pix_loss, 2.0f * st.inv_n,
                 st.partials, ext_grad, fwd_only, st.v, d->inv_mvp, st.bounds, d->alpha, st.galpha,
                 d->alpha_level, static_cast<int>(L.Sa)};
    hipError_t e;
    DS_TRY(launch_shade(L.nl, d->channels, d->hidden, L.sh, d->bf16 != 0, sa, L.nblocks, st.s, &st.va), "shade");
    return DS_OK;
}

// Edge gradients, opacity gradient, the closing reductions and (accumulate)
// the backward of the accumulations: everything after the shading adjoint.
int enqueue_backward(const ds_step_desc* d, Step& st, const float* grad_color, bool write_loss) {
    const StepLayout& L = st.L;
    const hipStream_t s = st.s;
    const bool acc = d->accumulate != 0;
    hipError_t e;
    DS_TRY(mark(d, s, DS_STAGE_EDGE, true), "event");
    DS_TRY(launch_edge_grad_sorted(st.v, 3, d->color, grad_color, d->rast, d->V, d->pos, st.gpos, d->T, d->tri,
                                   st.rw.recs, st.px_list, st.tile_px_off + L.n_tiles, L.P, s),
           "edge grad");
    DS_TRY(mark(d, s, DS_STAGE_EDGE, false), "event");
    DS_TRY(mark(d, s, DS_STAGE_AUX, true), "event");
    AlphaSource a{d->alpha, d->alpha_level, static_cast<int>(L.Sa)};
    DS_TRY(launch_opacity_aux_records(st.v, d->T, a, d->rast, d->pix_loss, st.galpha, st.rw, st.aux_rec,
                                      st.aux_n, L.aux_per_quad, st.va, s),
           "opacity aux");
    DS_TRY(mark(d, s, DS_STAGE_AUX, false), "event");
    DS_TRY(mark(d, s, DS_STAGE_FINAL, true), "event");
    {
        const long long n_alpha = acc ? 0 : (long long)d->T * L.Sa;   // accumulate: the sigmoid backward follows
        const long long nf = std::max<long long>(d->V, n_alpha);
        const unsigned R = static_cast<unsigned>((L.nparam + 63) / 64);
        k_epilogue<<<R + 2 + blocks_for(nf, 1024), 1024, 0, s>>>(
            st.partials, L.nblocks, L.nparam, d->grad_mlp, st.bgloss, L.n_tiles, st.inv_n,
            write_loss ? d->loss : nullptr, d->B, d->V, d->mvp, reinterpret_cast<const float4*>(st.gpos),
            d->grad_verts, n_alpha, d->alpha, st.galpha, d->grad_alpha_logits, st.rw.tile_cov, L.n_tiles,
            st.rw.stats,
            d->seed_device && d->advance_seed
                ? const_cast<unsigned long long*>(reinterpret_cast<const unsigned long long*>(d->seed_device))
                : nullptr,
            reinterpret_cast<long long*>(d->advance_counter));
        DS_TRY(hipGetLastError(), "epilogue");
        if (acc && d->T > 0) {
            DS_TRY(accum_lds_opt_in(), "accumulate LDS");
            k_accum_sigmoid_bwd<<<std::min(d->T, 4096), 256, accum_lds_bytes(d, L), s>>>(
                d->T, feat_job(d, st), alpha_job(d, st), accum_merge(d, L) ? 1 : 0);
            DS_TRY(hipGetLastError(), "accumulate + sigmoid backward");
        }
    }
    DS_TRY(mark(d, s, DS_STAGE_FINAL, false), "event");
    return DS_OK;
}

#undef DS_TRY

}  // namespace
}  // namespace ds

extern "C" {

int64_t ds_step_mlp_params(int levels, int channels, int hidden) {
    return ds::shade_nparam(levels, channels, hidden, false);
}

int64_t ds_step_mlp_params_sh2(int levels, int channels, int hidden, int sh2) {
    return ds::shade_nparam(levels, channels, hidden, sh2 != 0);
}

size_t ds_step_workspace_size(const ds_step_desc* d) {
    ds::Views v;
    if (!d || ds::step_make_views(d, v) != DS_OK) return 0;
    ds::StepLayout L;
    if (!ds::layout(d, v, L)) return 0;
    return L.total;
}

int ds_train_step(const ds_step_desc* d, void* stream) {
    using namespace ds;
    Step st;
    int rc = step_prepare(d, stream, st);
    if (rc != DS_OK) return rc;
    if ((rc = enqueue_forward(d, st, true)) != DS_OK) return rc;
    hipError_t e;
    if ((e = mark(d, st.s, DS_STAGE_SHADE, true)) != hipSuccess) return step_fail(DS_ERR_HIP, "event");
    if (d->loss_kind == DS_LOSS_L1_SSIM) {
        // the loss couples neighbouring pixels: colour first, then the
        // L1 + SSIM gradient of the whole image, then the adjoint on it
        if ((rc = enqueue_shade(d, st, 1, nullptr)) != DS_OK) return rc;
        if ((e = launch_photo_loss(st.v, 3, d->target, d->color, st.photo, 0.8f, 0.2f, d->grad_color, d->loss,
                                   st.s)) != hipSuccess)
            return step_fail(DS_ERR_HIP, "L1 + SSIM loss");
        if ((rc = enqueue_shade(d, st, 0, d->grad_color)) != DS_OK) return rc;
    } else {
        if ((rc = enqueue_shade(d, st, 0, nullptr)) != DS_OK) return rc;
    }
    if ((e = mark(d, st.s, DS_STAGE_SHADE, false)) != hipSuccess) return step_fail(DS_ERR_HIP, "event");
    return enqueue_backward(d, st, d->grad_color, d->loss_kind != DS_LOSS_L1_SSIM);
}

int ds_step_forward(const ds_step_desc* d, void* stream) {
    using namespace ds;
    Step st;
    int rc = step_prepare(d, stream, st);
    if (rc != DS_OK) return rc;
    if ((rc = enqueue_forward(d, st, false)) != DS_OK) return rc;
    return enqueue_shade(d, st, 1, nullptr);
}

int ds_step_backward(const ds_step_desc* d, const float* grad_color, void* stream) {
    using namespace ds;
    if (!grad_color) return step_fail(DS_ERR_INVALID, "null grad_color");
    if (reinterpret_cast<uintptr_t>(grad_color) & 3u) return step_fail(DS_ERR_INVALID, "grad_color misaligned");
    Step st;
    int rc = step_prepare(d, stream, st);
    if (rc != DS_OK) return rc;
    // the forward's ranges again, now clearing the rows the adjoint adds into
    if (launch_shade_ranges(shade_ranges(d, st, true), st.s) != hipSuccess)
        return step_fail(DS_ERR_HIP, "shading ranges");
    if ((rc = enqueue_shade(d, st, 0, grad_color)) != DS_OK) return rc;
    return enqueue_backward(d, st, grad_color, false);
}

int ds_autotune_step(ds_step_desc* d, const int* tiles, int n_tiles, int reps, float* ms, void* stream) {
    using namespace ds;
    if (!d || !tiles || !ms || n_tiles < 1 || n_tiles > 8 || reps < 1)
        return step_fail(DS_ERR_INVALID, "autotune: tiles[1..8], reps >= 1 and non-null outputs required");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int saved_stage = d->timed_stage;
    const int orig_tile = d->tile;
    d->timed_stage = 0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        if (e0) (void)hipEventDestroy(e0);
        d->timed_stage = saved_stage;
        return step_fail(DS_ERR_HIP, "autotune: event creation");
    }
    int rc = DS_OK, best = -1;
    for (int i = 0; i < n_tiles && rc == DS_OK; ++i) {
        ms[i] = -1.f;
        d->tile = tiles[i];
        const size_t need = ds_step_workspace_size(d);
        if (need == 0 || need > d->workspace_bytes) continue;   // candidate does not fit: skipped
        rc = ds_train_step(d, stream);                         // warm-up (code objects, caches)
        if (rc != DS_OK) break;
        if (hipEventRecord(e0, s) != hipSuccess) { rc = step_fail(DS_ERR_HIP, "autotune: event"); break; }
        for (int r = 0; r < reps && rc == DS_OK; ++r) rc = ds_train_step(d, stream);
        if (rc != DS_OK) break;
        float t = 0.f;
        if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&t, e0, e1) != hipSuccess) {
            rc = step_fail(DS_ERR_HIP, "autotune: timing");
            break;
        }
        ms[i] = t / static_cast<float>(reps);
        if (best < 0 || ms[i] < ms[best]) best = i;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    d->timed_stage = saved_stage;
    if (rc != DS_OK) {
        d->tile = orig_tile;
        return rc;
    }
    if (best < 0) {
        d->tile = orig_tile;
        return step_fail(DS_ERR_WORKSPACE, "autotune: no candidate tile fits the workspace");
    }
    d->tile = tiles[best];
    return DS_OK;
}

}  // extern "C"
