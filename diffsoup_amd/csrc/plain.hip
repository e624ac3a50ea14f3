// The plain FP32 path: the reference's rasterization structure implemented
// directly in HIP, kept as the comparator for the tile-binned path and as a
// one-to-one stand-in for each compiled function the reference's Python layer
// binds (src/main.cpp:30-108, 134-210):
//
//   rects      per (view, triangle) conservative rectangle and fragment count,
//              inclusive scan, total returned to the host
//   fragments  one thread per candidate: binary search for its triangle,
//              coverage, barycentrics and depth range (invalid -> view = -1)
//   compact    order-preserving compaction (a scan of the valid flags), so the
//              fragment list is in (view, triangle, rectangle position) order
//   alpha      per-fragment lattice opacity; thresholds from the mask hash
//   depth      global 64-bit atomicMin of bits(depth + 2) << 32 | fragment
//              index into a per-pixel key buffer, then a gather into rast
//   aux        per-fragment score-function opacity gradient and its scatter
//
// The host-visible counts (candidates, valid fragments) are returned with a
// blocking copy, as the reference does; everything else stays on the stream.
#include "diffsoup_hip.h"
#include "ds_kernels.h"
#include "ds_math.h"
#include "ds_scan.h"

namespace ds {
namespace {

constexpr int kBS = 256;
inline unsigned nblk(long long n) { return static_cast<unsigned>((n + kBS - 1) / kBS); }

__global__ void k_rects(int H, int W, int B, int V, const float4* __restrict__ pos, int T,
                        const int* __restrict__ tri, int4* __restrict__ rects, int* __restrict__ counts) {
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= (long long)B * T) return;
    const int b = static_cast<int>(i / T), t = static_cast<int>(i - (long long)b * T);
    const int i0 = tri[3 * t], i1 = tri[3 * t + 1], i2 = tri[3 * t + 2];
    int4 out = make_int4(0, 0, 0, 0);
    int cnt = 0;
    if (i0 >= 0 && i0 < V && i1 >= 0 && i1 < V && i2 >= 0 && i2 < V) {
        const float4* P = pos + (long long)b * V;
        const Rect r = triangle_rect(P[i0], P[i1], P[i2], H, W);
        if (r.y1 > r.y0) {
            out = make_int4(r.y0, r.y1 - r.y0, r.x0, r.x1 - r.x0);   // h0, h_len, w0, w_len
            cnt = (r.y1 - r.y0) * (r.x1 - r.x0);
        }
    }
    rects[i] = out;
    counts[i] = cnt;
}

__global__ void k_fragments(int H, int W, int V, const float4* __restrict__ pos, int T,
                            const int* __restrict__ tri, int num_tris, long long num_frags,
                            const long long* __restrict__ prefix, const int4* __restrict__ rects,
                            int* __restrict__ frag_pix, float4* __restrict__ frag_attrs) {
    const long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (idx >= num_frags) return;
    frag_pix[3 * idx] = -1;
    int lo = 0, hi = num_tris;   // first i with prefix[i] > idx (inclusive prefix)
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (prefix[mid] <= idx) lo = mid + 1; else hi = mid;
    }
    const int i = lo;
    const int b = i / T, t = i - b * T;
    const int4 r = rects[i];
    if (r.y <= 0 || r.w <= 0) return;
    const long long local = idx - (i == 0 ? 0 : prefix[i - 1]);
    const int y = r.x + static_cast<int>(local / r.w);
    const int x = r.z + static_cast<int>(local % r.w);
    if (x < 0 || x >= W || y < 0 || y >= H) return;
    const float4* P = pos + (long long)b * V;
    Hit h;
    if (!cover(pixel_ndc(y, H), pixel_ndc(x, W), P[tri[3 * t]], P[tri[3 * t + 1]], P[tri[3 * t + 2]], h)) return;
    float depth;
    if (!fragment_depth(h, depth)) return;
    frag_pix[3 * idx] = b;
    frag_pix[3 * idx + 1] = y;
    frag_pix[3 * idx + 2] = x;
    frag_attrs[idx] = make_float4(h.b0, h.b1, depth, static_cast<float>(t + 1));
}

__global__ void k_valid_flags(long long n, const int* __restrict__ frag_pix, int* __restrict__ flags) {
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i < n) flags[i] = frag_pix[3 * i] >= 0 ? 1 : 0;
}

__global__ void k_compact(long long n, const int* __restrict__ frag_pix, const float4* __restrict__ frag_attrs,
                          const int* __restrict__ pos_excl, int* __restrict__ pix_out,
                          float4* __restrict__ attrs_out) {
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= n || frag_pix[3 * i] < 0) return;
    const long long o = pos_excl[i];
    pix_out[3 * o] = frag_pix[3 * i];
    pix_out[3 * o + 1] = frag_pix[3 * i + 1];
    pix_out[3 * o + 2] = frag_pix[3 * i + 2];
    attrs_out[o] = frag_attrs[i];
}

// Per-fragment opacity summed over levels lo..hi (feature_dim 1), levels
// ascending and corners 0, 1, 2 from +0 as in the reference's interpolation
// (src/cuda/multires.cu:10-68, :130-156).
__global__ void k_alpha(long long n, const float4* __restrict__ attrs, int lo, int hi, int S,
                        const float* __restrict__ alpha_src, float* __restrict__ out) {
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 a = attrs[i];
    const int t = static_cast<int>(a.w) - 1;
    if (t < 0) { out[i] = 0.f; return; }
    const float* src = alpha_src + (long long)t * S;
    float v = 0.f;
    for (int l = lo; l <= hi; ++l) {
        const Corners c = lattice_corners(a.x, a.y, static_cast<uint32_t>(l));
        v += c.w[0] * src[c.idx[0]];
        v += c.w[1] * src[c.idx[1]];
        v += c.w[2] * src[c.idx[2]];
        src += feats_at_level(static_cast<uint32_t>(l));
    }
    out[i] = v;
}

__global__ void k_alpha_bwd(long long n, const float4* __restrict__ attrs, int lo, int hi, int S,
                            float* __restrict__ grad_src, const float* __restrict__ grad) {
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 a = attrs[i];
    const int t = static_cast<int>(a.w) - 1;
    if (t < 0) return;
    float* d = grad_src + (long long)t * S;
    const float g = grad[i];
    for (int l = lo; l <= hi; ++l) {
        const Corners c = lattice_corners(a.x, a.y, static_cast<uint32_t>(l));
        for (int k = 0; k < 3; ++k) unsafeAtomicAdd(&d[c.idx[k]], g * c.w[k]);
        d += feats_at_level(static_cast<uint32_t>(l));
    }
}

__global__ void k_thresholds(long long n, const int* __restrict__ frag_pix, const float4* __restrict__ attrs,
                             int stochastic, unsigned long long seed, int view_base, float* __restrict__ out) {
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = stochastic ? mask_threshold(seed, static_cast<uint32_t>(frag_pix[3 * i] + view_base),
                                         static_cast<uint32_t>(static_cast<int>(attrs[i].w) - 1),
                                         static_cast<uint32_t>(frag_pix[3 * i + 1]),
                                         static_cast<uint32_t>(frag_pix[3 * i + 2]))
                        : 0.5f;
}

__global__ void k_fill_keys(long long n, long long* __restrict__ z) {
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i < n) z[i] = kKeyEmpty;
}

__global__ void k_depth_atomic(int H, int W, long long n, const int* __restrict__ frag_pix,
                               const float4* __restrict__ attrs, const float* __restrict__ alpha,
                               const float* __restrict__ thresh, long long* __restrict__ z) {
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int b = frag_pix[3 * i];
    if (b < 0) return;
    if (alpha[i] < thresh[i]) return;
    const float d = attrs[i].z;
    if (!depth_testable(d)) return;
    const long long px = ((long long)b * H + frag_pix[3 * i + 1]) * W + frag_pix[3 * i + 2];
    atomicMin(&z[px], static_cast<long long>(depth_key(d, static_cast<uint32_t>(i))));
}

__global__ void k_depth_gather(long long P, const long long* __restrict__ z, const float4* __restrict__ attrs,
                               float4* __restrict__ rast) {
    const long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (p >= P) return;
    const long long k = z[p];
    rast[p] = k == kKeyEmpty ? make_float4(0.f, 0.f, 0.f, 0.f) : attrs[static_cast<uint32_t>(k & 0xffffffffll)];
}

__global__ void k_aux(int H, int W, int C, const float* __restrict__ color, const float* __restrict__ target,
                      const float4* __restrict__ rast, long long n, const int* __restrict__ frag_pix,
                      const float4* __restrict__ attrs, const float* __restrict__ alpha, float* __restrict__ grad) {
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int b = frag_pix[3 * i];
    if (b < 0) return;
    const long long p = ((long long)b * H + frag_pix[3 * i + 1]) * W + frag_pix[3 * i + 2];
    const float4 a = attrs[i];
    const float4 r = rast[p];
    const int tp = static_cast<int>(r.w) - 1;
    if (tp >= 0 && a.z > r.z) return;
    float l = 0.f;
    for (int c = 0; c < C; ++c) l += fabsf(color[p * C + c] - target[p * C + c]);
    constexpr float eps = 1e-5f;
    const float al = alpha[i];
    if (tp >= 0 && static_cast<int>(a.w) - 1 == tp) {
        grad[i] += l / (al > eps ? al : eps);
    } else {
        const float om = 1.f - al;
        grad[i] -= l / (om > eps ? om : eps);
    }
}

}  // namespace

int plain_fail(int code, const char* msg);

}  // namespace ds

extern "C" {

size_t ds_plain_scan_tmp_bytes(int64_t n) { return ds::scan_tmp_bytes(n); }

int64_t ds_plain_triangle_rects(int H, int W, int B, int V, const float* pos, int T, const int* tri,
                                int* rects, int64_t* prefix, int* counts, void* scan_tmp, void* stream) {
    using namespace ds;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (H < 1 || W < 1 || B < 1 || V < 0 || T < 0) return plain_fail(DS_ERR_INVALID, "bad sizes");
    const long long n = (long long)B * T;
    if (n == 0) return 0;
    if (!pos || !tri || !rects || !prefix || !counts || !scan_tmp) return plain_fail(DS_ERR_INVALID, "null pointer");
    k_rects<<<nblk(n), kBS, 0, s>>>(H, W, B, V, reinterpret_cast<const float4*>(pos), T, tri,
                                    reinterpret_cast<int4*>(rects), counts);
    long long total = 0;
    if (launch_inclusive_scan_i64(counts, n, reinterpret_cast<long long*>(prefix), scan_tmp, s) != hipSuccess ||
        hipMemcpyAsync(&total, scan_total_slot(scan_tmp, n), sizeof(long long), hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return plain_fail(DS_ERR_HIP, "rect scan");
    return total;
}

int ds_plain_fragments(int H, int W, int B, int V, const float* pos, int T, const int* tri, const int64_t* prefix,
                       const int* rects, int64_t num_frags, int* frag_pix, float* frag_attrs, void* stream) {
    using namespace ds;
    if (num_frags <= 0) return DS_OK;
    if (num_frags > INT32_MAX) return plain_fail(DS_ERR_INVALID, "more than 2^31 candidate fragments");
    k_fragments<<<nblk(num_frags), kBS, 0, static_cast<hipStream_t>(stream)>>>(
        H, W, V, reinterpret_cast<const float4*>(pos), T, tri, B * T, num_frags,
        reinterpret_cast<const long long*>(prefix), reinterpret_cast<const int4*>(rects), frag_pix,
        reinterpret_cast<float4*>(frag_attrs));
    return hipGetLastError() == hipSuccess ? DS_OK : plain_fail(DS_ERR_HIP, "fragments");
}

int64_t ds_plain_compact(int64_t num_frags, const int* frag_pix, const float* frag_attrs, int* frag_pix_out,
                         float* frag_attrs_out, int* flags, void* scan_tmp, void* stream) {
    using namespace ds;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (num_frags <= 0) return 0;
    k_valid_flags<<<nblk(num_frags), kBS, 0, s>>>(num_frags, frag_pix, flags);
    long long total = 0;
    // exclusive positions overwrite the flags in place
    if (launch_exclusive_scan_inplace(flags, num_frags, scan_tmp, s) != hipSuccess ||
        hipMemcpyAsync(&total, scan_total_slot(scan_tmp, num_frags), sizeof(long long), hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return plain_fail(DS_ERR_HIP, "compaction scan");
    k_compact<<<nblk(num_frags), kBS, 0, s>>>(num_frags, frag_pix, reinterpret_cast<const float4*>(frag_attrs),
                                              flags, frag_pix_out, reinterpret_cast<float4*>(frag_attrs_out));
    return hipGetLastError() == hipSuccess ? total : plain_fail(DS_ERR_HIP, "compact");
}

int ds_plain_alpha(int64_t n, const float* frag_attrs, int min_level, int max_level, const float* alpha_src,
                   float* frag_alpha, void* stream) {
    using namespace ds;
    if (min_level < 0 || max_level < min_level || max_level > 10)
        return plain_fail(DS_ERR_INVALID, "alpha: levels must satisfy 0 <= min_level <= max_level <= 10");
    if (n <= 0) return DS_OK;
    k_alpha<<<nblk(n), kBS, 0, static_cast<hipStream_t>(stream)>>>(
        n, reinterpret_cast<const float4*>(frag_attrs), min_level, max_level,
        static_cast<int>(feats_in_levels(min_level, max_level)), alpha_src, frag_alpha);
    return hipGetLastError() == hipSuccess ? DS_OK : plain_fail(DS_ERR_HIP, "alpha");
}

int ds_plain_alpha_bwd(int64_t n, const float* frag_attrs, int min_level, int max_level, float* grad_alpha_src,
                       const float* grad_frag_alpha, void* stream) {
    using namespace ds;
    if (min_level < 0 || max_level < min_level || max_level > 10)
        return plain_fail(DS_ERR_INVALID, "alpha bwd: levels must satisfy 0 <= min_level <= max_level <= 10");
    if (n <= 0) return DS_OK;
    k_alpha_bwd<<<nblk(n), kBS, 0, static_cast<hipStream_t>(stream)>>>(
        n, reinterpret_cast<const float4*>(frag_attrs), min_level, max_level,
        static_cast<int>(feats_in_levels(min_level, max_level)), grad_alpha_src, grad_frag_alpha);
    return hipGetLastError() == hipSuccess ? DS_OK : plain_fail(DS_ERR_HIP, "alpha bwd");
}

int ds_plain_thresholds(int64_t n, const int* frag_pix, const float* frag_attrs, int stochastic, uint64_t seed,
                        int view_base, float* thresh, void* stream) {
    using namespace ds;
    if (n <= 0) return DS_OK;
    k_thresholds<<<nblk(n), kBS, 0, static_cast<hipStream_t>(stream)>>>(
        n, frag_pix, reinterpret_cast<const float4*>(frag_attrs), stochastic, seed, view_base, thresh);
    return hipGetLastError() == hipSuccess ? DS_OK : plain_fail(DS_ERR_HIP, "thresholds");
}

int ds_plain_depth_test(int B, int H, int W, int64_t n, const int* frag_pix, const float* frag_attrs,
                        const float* frag_alpha, const float* thresh, int64_t* zbuf, float* rast, void* stream) {
    using namespace ds;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const long long P = (long long)B * H * W;
    if (P <= 0) return DS_OK;
    long long* z = reinterpret_cast<long long*>(zbuf);
    k_fill_keys<<<nblk(P), kBS, 0, s>>>(P, z);
    if (n > 0)
        k_depth_atomic<<<nblk(n), kBS, 0, s>>>(H, W, n, frag_pix, reinterpret_cast<const float4*>(frag_attrs),
                                               frag_alpha, thresh, z);
    k_depth_gather<<<nblk(P), kBS, 0, s>>>(P, z, reinterpret_cast<const float4*>(frag_attrs),
                                           reinterpret_cast<float4*>(rast));
    return hipGetLastError() == hipSuccess ? DS_OK : plain_fail(DS_ERR_HIP, "depth test");
}

int ds_plain_opacity_aux(int B, int H, int W, int C, const float* color, const float* target, const float* rast,
                         int64_t n, const int* frag_pix, const float* frag_attrs, const float* frag_alpha,
                         float* grad_frag_alpha, void* stream) {
    using namespace ds;
    (void)B;
    if (n <= 0) return DS_OK;
    k_aux<<<nblk(n), kBS, 0, static_cast<hipStream_t>(stream)>>>(
        H, W, C, color, target, reinterpret_cast<const float4*>(rast), n, frag_pix,
        reinterpret_cast<const float4*>(frag_attrs), frag_alpha, grad_frag_alpha);
    return hipGetLastError() == hipSuccess ? DS_OK : plain_fail(DS_ERR_HIP, "opacity aux");
}

}  // extern "C"
