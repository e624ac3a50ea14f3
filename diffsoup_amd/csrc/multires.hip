// Multires barycentric-lattice features: per-pixel gather (sum of levels as
// in the reference, or per-level concatenation), its scatter adjoint, and the
// cross-level accumulation to one target lattice with its adjoint.
//
// Reference: src/cuda/multires.cu:10-128 (interp / scatter device
// functions), :224-312 (per-pixel colour), :313-450 (accumulate_to_level).
// Per output element the summation order is the reference's (levels
// ascending, corners 0, 1, 2, starting from +0), so the sum-mode gather is
// bit-identical to it.
#include "ds_kernels.h"
#include "ds_math.h"

namespace ds {
namespace {

constexpr int kMaxLevels = 11;  // levels 0..10

// Per-pixel gather.  One thread per pixel; background pixels get zeros.
// Layout: rast [P, 4]; feat [T, S_tot, C]; out [P, C] (sum) or
// [P, (hi - lo + 1) * C] (concat, level-major then channel).
__global__ void k_color_fwd(long long P, const float4* __restrict__ rast, int lo, int hi, int T, int C,
                            int S_tot, const float* __restrict__ feat, int concat,
                            float* __restrict__ out) {
    const long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (p >= P) return;
    const int nl = hi - lo + 1;
    const int width = concat ? nl * C : C;
    float* o = out + p * width;
    const float4 r = rast[p];
    const int t = static_cast<int>(r.w) - 1;
    if (t < 0 || t >= T) {
        for (int j = 0; j < width; ++j) o[j] = 0.f;
        return;
    }
    Corners cs[kMaxLevels];
    uint32_t off[kMaxLevels];
    uint32_t acc = 0;
    for (int l = lo; l <= hi; ++l) {
        cs[l - lo] = lattice_corners(r.x, r.y, static_cast<uint32_t>(l));
        off[l - lo] = acc;
        acc += feats_at_level(static_cast<uint32_t>(l));
    }
    const float* f = feat + (long long)t * S_tot * C;
    if (concat) {
        for (int k = 0; k < nl; ++k) {
            const float* fl = f + (long long)off[k] * C;
            for (int j = 0; j < C; ++j) {
                float v = 0.f;
                v += cs[k].w[0] * fl[cs[k].idx[0] * C + j];
                v += cs[k].w[1] * fl[cs[k].idx[1] * C + j];
                v += cs[k].w[2] * fl[cs[k].idx[2] * C + j];
                o[k * C + j] = v;
            }
        }
    } else {
        for (int j = 0; j < C; ++j) {
            float v = 0.f;
            for (int k = 0; k < nl; ++k) {
                const float* fl = f + (long long)off[k] * C;
                v += cs[k].w[0] * fl[cs[k].idx[0] * C + j];
                v += cs[k].w[1] * fl[cs[k].idx[1] * C + j];
                v += cs[k].w[2] * fl[cs[k].idx[2] * C + j];
            }
            o[j] = v;
        }
    }
}

__global__ void k_color_bwd(long long P, const float4* __restrict__ rast, int lo, int hi, int T, int C,
                            int S_tot, float* __restrict__ grad_feat, int concat,
                            const float* __restrict__ grad_out) {
    const long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (p >= P) return;
    const float4 r = rast[p];
    const int t = static_cast<int>(r.w) - 1;
    if (t < 0 || t >= T) return;
    const int nl = hi - lo + 1;
    const float* g = grad_out + p * (concat ? nl * C : C);
    float* gf = grad_feat + (long long)t * S_tot * C;
    uint32_t off = 0;
    for (int l = lo; l <= hi; ++l) {
        const Corners c = lattice_corners(r.x, r.y, static_cast<uint32_t>(l));
        const float* gl = concat ? g + (l - lo) * C : g;
        float* dst = gf + (long long)off * C;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < C; ++j) unsafeAtomicAdd(&dst[c.idx[i] * C + j], gl[j] * c.w[i]);
        off += feats_at_level(static_cast<uint32_t>(l));
    }
}


__global__ void k_accum_fwd(int T, int lo, int hi, int target, int S_tot, int S_T, int C,
                            const float* __restrict__ feat, float* __restrict__ out) {
    const long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (id >= (long long)T * S_T) return;
    const int t = static_cast<int>(id / S_T);
    const uint32_t k = static_cast<uint32_t>(id - (long long)t * S_T);
    uint32_t x, y;
    lattice_xy(k, x, y);
    const float Rf = static_cast<float>(1u << target);
    const float b0 = static_cast<float>(x) / Rf, b1 = static_cast<float>(y) / Rf;
    const float* f = feat + (long long)t * S_tot * C;
    float* o = out + id * C;
    for (int j = 0; j < C; ++j) {
        float v = 0.f;
        uint32_t off = 0;
        for (int l = lo; l <= hi; ++l) {
            const Corners c = lattice_corners(b0, b1, static_cast<uint32_t>(l));
            const float* fl = f + (long long)off * C;
            v += c.w[0] * fl[c.idx[0] * C + j];
            v += c.w[1] * fl[c.idx[1] * C + j];
            v += c.w[2] * fl[c.idx[2] * C + j];
            off += feats_at_level(static_cast<uint32_t>(l));
        }
        o[j] = v;
    }
}

__global__ void k_accum_bwd(int T, int lo, int hi, int target, int S_tot, int S_T, int C,
                            float* __restrict__ grad_feat, const float* __restrict__ grad_out) {
    const long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (id >= (long long)T * S_T) return;
    const int t = static_cast<int>(id / S_T);
    const uint32_t k = static_cast<uint32_t>(id - (long long)t * S_T);
    uint32_t x, y;
    lattice_xy(k, x, y);
    const float Rf = static_cast<float>(1u << target);
    const float b0 = static_cast<float>(x) / Rf, b1 = static_cast<float>(y) / Rf;
    float* gf = grad_feat + (long long)t * S_tot * C;
    const float* g = grad_out + id * C;
    uint32_t off = 0;
    for (int l = lo; l <= hi; ++l) {
        const Corners c = lattice_corners(b0, b1, static_cast<uint32_t>(l));
        float* dst = gf + (long long)off * C;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < C; ++j) unsafeAtomicAdd(&dst[c.idx[i] * C + j], g[j] * c.w[i]);
        off += feats_at_level(static_cast<uint32_t>(l));
    }
}

inline unsigned blocks_for(long long n, int bs) { return static_cast<unsigned>((n + bs - 1) / bs); }

}  // namespace

hipError_t launch_multires_color_fwd(long long P, const float* rast, int lo, int hi, int T, int C,
                                     const float* feat, int concat, float* out, hipStream_t s) {
    if (P <= 0) return hipSuccess;
    const int S_tot = static_cast<int>(feats_in_levels(lo, hi));
    k_color_fwd<<<blocks_for(P, 256), 256, 0, s>>>(P, reinterpret_cast<const float4*>(rast), lo, hi, T, C,
                                                   S_tot, feat, concat, out);
    return hipGetLastError();
}

hipError_t launch_multires_color_bwd(long long P, const float* rast, int lo, int hi, int T, int C,
                                     float* grad_feat, int concat, const float* grad_out,
                                     hipStream_t s) {
    if (P <= 0) return hipSuccess;
    const int S_tot = static_cast<int>(feats_in_levels(lo, hi));
    k_color_bwd<<<blocks_for(P, 256), 256, 0, s>>>(P, reinterpret_cast<const float4*>(rast), lo, hi, T, C,
                                                   S_tot, grad_feat, concat, grad_out);
    return hipGetLastError();
}

hipError_t launch_accumulate_fwd(int T, int lo, int hi, int target, int C, const float* feat,
                                 float* out, hipStream_t s) {
    const int S_tot = static_cast<int>(feats_in_levels(lo, hi));
    const int S_T = static_cast<int>(feats_at_level(target));
    const long long n = (long long)T * S_T;
    if (n <= 0) return hipSuccess;
    k_accum_fwd<<<blocks_for(n, 256), 256, 0, s>>>(T, lo, hi, target, S_tot, S_T, C, feat, out);
    return hipGetLastError();
}

hipError_t launch_accumulate_bwd(int T, int lo, int hi, int target, int C, float* grad_feat,
                                 const float* grad_out, hipStream_t s) {
    const int S_tot = static_cast<int>(feats_in_levels(lo, hi));
    const int S_T = static_cast<int>(feats_at_level(target));
    const long long n = (long long)T * S_T;
    if (n <= 0) return hipSuccess;
    k_accum_bwd<<<blocks_for(n, 256), 256, 0, s>>>(T, lo, hi, target, S_tot, S_T, C, grad_feat,
                                                   grad_out);
    return hipGetLastError();
}

}  // namespace ds
