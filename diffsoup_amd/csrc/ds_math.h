// Scalar math shared by every DiffSoup HIP kernel: pixel centres, the
// conservative screen rectangle, perspective-correct coverage, the packed
// depth key, the barycentric lattice and the stochastic-mask hash.
//
// Every float expression here keeps the operation order of the reference
// kernels and the whole library is compiled with -ffp-contract=off, so a
// multiply followed by an add is never fused.  Together with IEEE division
// (hipcc's default correctly rounded fp32 divide) this makes the coverage,
// barycentric and depth values bit-identical to a plain C evaluation of the
// same expressions, which is what the parity tests demand.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ds {

// ---------------------------------------------------------------------------
// Pixel centre in NDC: -1 + 2 (i + 0.5) / n.  Row 0 maps to ndc_y = -1.
// Reference: src/cuda/rasterize.cu:250-251.
// ---------------------------------------------------------------------------
__host__ __device__ __forceinline__ float pixel_ndc(int i, int n) {
    return -1.f + 2.f * (static_cast<float>(i) + 0.5f) / static_cast<float>(n);
}

// ---------------------------------------------------------------------------
// Conservative pixel rectangle of a clip-space triangle, or an empty one when
// the triangle is culled.  Reference: src/cuda/rasterize.cu:10-137
// (w <= 0 cull :77, outcode reject :84-91, NDC clamp with non-finite -> +-1
// :18-41, floor-based pixel bounds :108-126).  The returned rectangle is
// half-open: rows [y0, y1), columns [x0, x1).
// ---------------------------------------------------------------------------
struct Rect { int y0, y1, x0, x1; };

__host__ __device__ __forceinline__ int outcode(float x, float y, float z, float w) {
    int c = 0;
    c |= (x < -w) ? 1 : 0;   c |= (x > w) ? 2 : 0;
    c |= (y < -w) ? 4 : 0;   c |= (y > w) ? 8 : 0;
    c |= (z < -w) ? 16 : 0;  c |= (z > w) ? 32 : 0;
    return c;
}

__host__ __device__ __forceinline__ float clamp_ndc(float v) {
    if (!isfinite(v)) v = (v > 0.f ? 1.f : -1.f);
    return fminf(1.f, fmaxf(-1.f, v));
}

__host__ __device__ __forceinline__ Rect triangle_rect(float4 p0, float4 p1, float4 p2, int H, int W) {
    Rect r{0, 0, 0, 0};
    if (!(p0.w > 0.f && p1.w > 0.f && p2.w > 0.f)) return r;
    if ((outcode(p0.x, p0.y, p0.z, p0.w) & outcode(p1.x, p1.y, p1.z, p1.w) &
         outcode(p2.x, p2.y, p2.z, p2.w)) != 0) return r;

    float xmin = 1.f, ymin = 1.f, zmin = 1.f, xmax = -1.f, ymax = -1.f, zmax = -1.f;
    const float4 pv[3] = {p0, p1, p2};
    for (int i = 0; i < 3; ++i) {
        // the reference multiplies by a rounded reciprocal, not a quotient
        const float iw = 1.f / pv[i].w;
        const float nx = clamp_ndc(pv[i].x * iw);
        const float ny = clamp_ndc(pv[i].y * iw);
        const float nz = clamp_ndc(pv[i].z * iw);
        xmin = fminf(xmin, nx); xmax = fmaxf(xmax, nx);
        ymin = fminf(ymin, ny); ymax = fmaxf(ymax, ny);
        zmin = fminf(zmin, nz); zmax = fmaxf(zmax, nz);
    }
    if (zmax < -1.f || zmin > 1.f) return r;

    const float fW = static_cast<float>(W), fH = static_cast<float>(H);
    float x0f = 0.5f * (xmin + 1.f) * fW, x1f = 0.5f * (xmax + 1.f) * fW;
    float y0f = 0.5f * (ymin + 1.f) * fH, y1f = 0.5f * (ymax + 1.f) * fH;
    if (x0f > x1f) { const float t = x0f; x0f = x1f; x1f = t; }
    if (y0f > y1f) { const float t = y0f; y0f = y1f; y1f = t; }

    const int x0 = max(static_cast<int>(floorf(x0f)), 0);
    const int x1 = min(static_cast<int>(floorf(x1f)), W - 1) + 1;
    const int y0 = max(static_cast<int>(floorf(y0f)), 0);
    const int y1 = min(static_cast<int>(floorf(y1f)), H - 1) + 1;
    if (x1 - x0 <= 0 || y1 - y0 <= 0) return r;
    r.y0 = y0; r.y1 = y1; r.x0 = x0; r.x1 = x1;
    return r;
}

// ---------------------------------------------------------------------------
// Perspective-correct coverage of one pixel centre by a clip-space triangle.
// Homogeneous 2D edge functions q_i = p_i.xy - ndc * p_i.w; b = A_i * (1/A);
// inclusive inside test; returns false for A == 0.
// Reference: src/cuda/rasterize.cu:170-206.
// ---------------------------------------------------------------------------
struct Hit { float b0, b1, z, w; };

__host__ __device__ __forceinline__ bool cover(float ndc_y, float ndc_x,
                                               float4 p0, float4 p1, float4 p2, Hit& h) {
    const float q0x = p0.x - ndc_x * p0.w, q0y = p0.y - ndc_y * p0.w;
    const float q1x = p1.x - ndc_x * p1.w, q1y = p1.y - ndc_y * p1.w;
    const float q2x = p2.x - ndc_x * p2.w, q2y = p2.y - ndc_y * p2.w;
    const float A0 = q1x * q2y - q1y * q2x;
    const float A1 = q2x * q0y - q2y * q0x;
    const float A2 = q0x * q1y - q0y * q1x;
    const float A = A0 + A1 + A2;
    if (A == 0.f) return false;
    // Early rejection that changes no decision: when A_i and A differ in sign
    // and |A_i| >= |A| 2^-60, b_i = A_i * (1/A) rounds to a non-zero negative
    // value (it cannot underflow to -0), so the inclusive test below fails;
    // the division is skipped for those pixels only.
    {
        const bool neg = A < 0.f;
        const float lim = fabsf(A) * 0x1p-60f;
        if ((A0 < 0.f) != neg && A0 != 0.f && fabsf(A0) >= lim) return false;
        if ((A1 < 0.f) != neg && A1 != 0.f && fabsf(A1) >= lim) return false;
    }
    const float invA = 1.f / A;
    h.b0 = A0 * invA;
    h.b1 = A1 * invA;
    const float b2 = 1.0f - h.b0 - h.b1;
    h.z = h.b0 * p0.z + h.b1 * p1.z + b2 * p2.z;
    h.w = h.b0 * p0.w + h.b1 * p1.w + b2 * p2.w;
    return h.b0 >= 0.f && h.b1 >= 0.f && h.b0 + h.b1 <= 1.f;
}

// Fragment depth z/w and the fragment-list range test [-1, 1]
// (src/cuda/rasterize.cu:266-268).
__host__ __device__ __forceinline__ bool fragment_depth(const Hit& h, float& depth) {
    depth = h.z / h.w;
    return depth >= -1.f && depth <= 1.f;
}

// Depth-test admission: depth strictly greater than -1 and finite
// (src/cuda/rasterize.cu:341).
__host__ __device__ __forceinline__ bool depth_testable(float depth) {
    return !(depth <= -1.f || !isfinite(depth));
}

// Packed key bits(depth + 2) << 32 | id, compared as a signed 64-bit integer;
// depth + 2 lies in (1, 3] so its bit pattern orders like its value.  The low
// word is the per-view triangle id, which reproduces the reference's
// smallest-fragment-index tie-break (src/cuda/rasterize.cu:308-312,343)
// because its fragment list is ordered by triangle within a view.
__host__ __device__ __forceinline__ int64_t depth_key(float depth, uint32_t id) {
    const float d2 = depth + 2.f;
    uint32_t bits;
    __builtin_memcpy(&bits, &d2, 4);
    return (static_cast<int64_t>(static_cast<int32_t>(bits)) << 32) | static_cast<int64_t>(id);
}
constexpr int64_t kKeyEmpty = INT64_MAX;

// ---------------------------------------------------------------------------
// Barycentric lattice at level L (R = 2^L subdivisions per edge).
// Sample (x, y) with x + y <= R lives at index (x+y)(x+y+1)/2 + y; a level
// stores (R+1)(R+2)/2 samples (3 at level 0).  A point (b0, b1) falls in a
// lattice cell whose three corners and weights are returned here.
// Reference: src/cuda/multires.cu:10-68, src/cuda/multires.cuh:8-20.
// ---------------------------------------------------------------------------
__host__ __device__ __forceinline__ uint32_t feats_at_level(uint32_t level) {
    return level == 0u ? 3u : (((1u << (level - 1u)) + 1u) * ((1u << level) + 1u));
}

__host__ __device__ __forceinline__ uint32_t feats_in_levels(uint32_t lo, uint32_t hi) {
    uint32_t s = 0;
    for (uint32_t l = lo; l <= hi; ++l) s += feats_at_level(l);
    return s;
}

struct Corners { uint32_t idx[3]; float w[3]; };

// Index of lattice sample (x, y).  Levels are at most 10 (R = 1024), so
// x + y < 2^23 and the product takes the 24-bit multiplier.
__host__ __device__ __forceinline__ uint32_t lattice_index(uint32_t x, uint32_t y) {
    const uint32_t n = (x + y) & 0x7fffffu;
    return ((n * (n + 1u)) >> 1) + y;
}

__host__ __device__ __forceinline__ Corners lattice_corners(float b0, float b1, uint32_t level) {
    const uint32_t R = 1u << level;
    const float Rf = static_cast<float>(R);
    float u = b0 * Rf, v = b1 * Rf;
    const uint32_t x = min(static_cast<uint32_t>(floorf(u)), R - 1u);
    const uint32_t y = min(static_cast<uint32_t>(floorf(v)), R - 1u - x);
    u = u - static_cast<float>(x);
    v = v - static_cast<float>(y);
    const bool flip = u + v > 1.f;
    const uint32_t fu = flip ? 1u : 0u;
    const float ff = flip ? 1.f : 0.f;
    const uint32_t x0 = x + 1u, y0 = y;
    const uint32_t x1 = x, y1 = y + 1u;
    const uint32_t x2 = x + fu, y2 = min(y + fu, R - x2);
    Corners c;
    c.idx[0] = lattice_index(x0, y0);
    c.idx[1] = lattice_index(x1, y1);
    c.idx[2] = lattice_index(x2, y2);
    c.w[0] = (1.f - ff) * u + ff * (1.f - v);
    c.w[1] = (1.f - ff) * v + ff * (1.f - u);
    c.w[2] = 1.f - c.w[0] - c.w[1];
    return c;
}

// Decode a lattice index into (x, y) with x + y = n, index = n(n+1)/2 + y
// (src/cuda/multires.cu:313-325).
__host__ __device__ __forceinline__ void lattice_xy(uint32_t k, uint32_t& x, uint32_t& y) {
    const float fi = static_cast<float>(k);
    const uint32_t n = static_cast<uint32_t>(floorf((sqrtf(8.f * fi + 1.f) - 1.f) * 0.5f));
    const uint32_t Tn = (n * (n + 1u)) >> 1;
    y = k - Tn;
    x = n - y;
}

// ---------------------------------------------------------------------------
// Stochastic-mask threshold: a counter-based hash of (seed, view, triangle,
// row, column) mapped to a 24-bit uniform in [0, 1).  A fragment is kept when
// alpha >= u (src/cuda/rasterize.cu:338 discards alpha < u).  The hash is
// SplitMix64's output function (Steele, Lea & Flood, OOPSLA 2014) applied
// twice; it replaces the reference's per-fragment torch.rand stream
// (python/diffsoup/multires.py:97-98), whose assignment to fragments depends
// on an atomic race, with a value that depends only on the fragment.
// ---------------------------------------------------------------------------
__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// The two rounds: the first depends on (seed, view) only, so a kernel that
// draws many thresholds of one view forms it once.
__host__ __device__ __forceinline__ uint64_t mask_view_hash(uint64_t seed, uint32_t view) {
    return mix64(seed + 0x9e3779b97f4a7c15ull * (static_cast<uint64_t>(view) + 1ull));
}

__host__ __device__ __forceinline__ float mask_threshold_in_view(uint64_t view_hash, uint32_t tri, uint32_t y,
                                                                 uint32_t x) {
    const uint64_t h = mix64(view_hash ^ ((static_cast<uint64_t>(tri) << 32) | (static_cast<uint64_t>(y) << 16) |
                                          static_cast<uint64_t>(x)));
    return static_cast<float>(static_cast<uint32_t>(h >> 40)) * (1.0f / 16777216.0f);
}

__host__ __device__ __forceinline__ float mask_threshold(uint64_t seed, uint32_t view, uint32_t tri,
                                                         uint32_t y, uint32_t x) {
    return mask_threshold_in_view(mask_view_hash(seed, view), tri, y, x);
}

// ---------------------------------------------------------------------------
// Degree-2 real spherical-harmonic basis of a pixel's view direction.  The
// direction points from the pixel ray's far-plane point to its near-plane
// point (toward the camera), both unprojected with the view's row-major
// inverse MVP; it is normalised with a correctly rounded 1 / sqrt.
// Reference: src/cuda/encoding.cu:9-85 (constants and sign pattern :56-85).
// ---------------------------------------------------------------------------
__host__ __device__ __forceinline__ void sh2_unproject(const float* m, float x, float y, float z, float out[3]) {
    const float cx = m[0] * x + m[1] * y + m[2] * z + m[3];
    const float cy = m[4] * x + m[5] * y + m[6] * z + m[7];
    const float cz = m[8] * x + m[9] * y + m[10] * z + m[11];
    const float cw = m[12] * x + m[13] * y + m[14] * z + m[15];
    const float iw = 1.0f / cw;
    out[0] = cx * iw; out[1] = cy * iw; out[2] = cz * iw;
}

__host__ __device__ __forceinline__ void sh2_basis(const float* inv_mvp, float ndc_x, float ndc_y, float e[9]) {
    // 1/(2 sqrt(pi)), sqrt(3)/(2 sqrt(pi)), sqrt(15)/(2 sqrt(pi)), sqrt(5)/(4 sqrt(pi)), sqrt(15)/(4 sqrt(pi))
    constexpr float kY00 = 0.28209479177387814f;
    constexpr float kY1 = 0.4886025119029199f;
    constexpr float kY2a = 1.0925484305920792f;
    constexpr float kY2b = 0.31539156525252005f;
    constexpr float kY2c = 0.5462742152960396f;
    float wn[3], wf[3];
    sh2_unproject(inv_mvp, ndc_x, ndc_y, -1.0f, wn);
    sh2_unproject(inv_mvp, ndc_x, ndc_y, 1.0f, wf);
    const float dx = wn[0] - wf[0], dy = wn[1] - wf[1], dz = wn[2] - wf[2];
    const float il = 1.0f / sqrtf(dx * dx + dy * dy + dz * dz);
    const float X = dx * il, Y = dy * il, Z = dz * il;
    const float xx = X * X, yy = Y * Y, zz = Z * Z;
    e[0] = kY00;
    e[1] = -kY1 * Y;
    e[2] = kY1 * Z;
    e[3] = -kY1 * X;
    e[4] = kY2a * (X * Y);
    e[5] = -kY2a * (Y * Z);
    e[6] = kY2b * (2.0f * zz - xx - yy);
    e[7] = -kY2a * (X * Z);
    e[8] = kY2c * (xx - yy);
}

}  // namespace ds
