// Image-space edge gradients w.r.t. clip-space vertex positions.
//
// For every covered pixel and each of its four neighbours (+x, -x, -y, +y)
// the colour difference weighted by the mean output gradient is an adjoint
// for moving the boundary between the two pixels.  Depending on what
// separates them it is routed to: the pixel's own triangle (interior motion
// or an opacity edge), a triangle-triangle intersection line, or the crossed
// geometric edge of the pixel's triangle.  This follows
// src/cuda/rasterize.cu:489-903 case for case, with correctly rounded
// 1/sqrt in place of the approximate reciprocal square root.
#include <algorithm>

#include "ds_kernels.h"
#include "ds_math.h"

namespace ds {
namespace {

// Segment (x0,y0)-(x1,y1) against (x2,y2)-(x3,y3); inclusive with tolerance,
// collinear overlaps report the middle of the overlap
// (src/cuda/rasterize.cu:493-544).
__device__ __forceinline__ bool segments_meet(float x0, float y0, float x1, float y1, float x2,
                                              float y2, float x3, float y3, float& ix, float& iy) {
    constexpr float eps = 1e-7f;
    const float rx = x1 - x0, ry = y1 - y0;
    const float sx = x3 - x2, sy = y3 - y2;
    const float den = rx * sy - ry * sx;
    const float qx = x2 - x0, qy = y2 - y0;
    if (fabsf(den) <= eps) {
        const float c = qx * ry - qy * rx;
        if (fabsf(c) > eps) return false;
        const float rr = rx * rx + ry * ry;
        if (rr < eps) return false;
        const float t0 = ((x2 - x0) * rx + (y2 - y0) * ry) / rr;
        const float t1 = ((x3 - x0) * rx + (y3 - y0) * ry) / rr;
        const float tmin = fminf(t0, t1), tmax = fmaxf(t0, t1);
        if (tmax < -eps || tmin > 1.f + eps) return false;
        const float t = fmaxf(0.f, fminf(1.f, 0.5f * (fmaxf(0.f, tmin) + fminf(1.f, tmax))));
        ix = x0 + t * rx;
        iy = y0 + t * ry;
        return true;
    }
    const float t = (qx * sy - qy * sx) / den;
    const float u = (qx * ry - qy * rx) / den;
    if (t < -eps || t > 1.f + eps || u < -eps || u > 1.f + eps) return false;
    ix = x0 + t * rx;
    iy = y0 + t * ry;
    return true;
}

// Which edge of the triangle the pixel-to-pixel segment crosses, and the
// perspective-correct position along it (src/cuda/rasterize.cu:546-604).
__device__ __forceinline__ bool crossed_edge(const float4 P[3], float x0, float y0, float x1, float y1,
                                             int& i0, int& i1, float& s0, float& s1) {
    float rx[3], ry[3];
    for (int k = 0; k < 3; ++k) { rx[k] = P[k].x / P[k].w; ry[k] = P[k].y / P[k].w; }
    float ix, iy;
    float4 A, B;
    if (segments_meet(x0, y0, x1, y1, rx[0], ry[0], rx[1], ry[1], ix, iy)) { i0 = 0; i1 = 1; A = P[0]; B = P[1]; }
    else if (segments_meet(x0, y0, x1, y1, rx[1], ry[1], rx[2], ry[2], ix, iy)) { i0 = 1; i1 = 2; A = P[1]; B = P[2]; }
    else if (segments_meet(x0, y0, x1, y1, rx[2], ry[2], rx[0], ry[0], ix, iy)) { i0 = 2; i1 = 0; A = P[2]; B = P[0]; }
    else return false;
    const float dx = B.x - A.x, dy = B.y - A.y, dw = B.w - A.w;
    const float ax = A.x - ix * A.w, ay = A.y - iy * A.w;
    const float bx = dx - ix * dw, by = dy - iy * dw;
    const float den = bx * bx + by * by;
    const float t = (den != 0.0f) ? -(ax * bx + ay * by) / den : 0.5f;
    s0 = 1.0f - t;
    s1 = t;
    return true;
}

// Unit normal of the triangle in NDC (src/cuda/rasterize.cu:608-633).
__device__ __forceinline__ bool ndc_normal(const float4 P[3], float n[3]) {
    if (fabsf(P[0].w) < 1e-30f || fabsf(P[1].w) < 1e-30f || fabsf(P[2].w) < 1e-30f) return false;
    float X[3], Y[3], Z[3];
    for (int k = 0; k < 3; ++k) { X[k] = P[k].x / P[k].w; Y[k] = P[k].y / P[k].w; Z[k] = P[k].z / P[k].w; }
    const float e1x = X[1] - X[0], e1y = Y[1] - Y[0], e1z = Z[1] - Z[0];
    const float e2x = X[2] - X[0], e2y = Y[2] - Y[0], e2z = Z[2] - Z[0];
    const float nx = e1y * e2z - e1z * e2y;
    const float ny = e1z * e2x - e1x * e2z;
    const float nz = e1x * e2y - e1y * e2x;
    const float l2 = nx * nx + ny * ny + nz * nz;
    if (!(l2 > 0.f)) return false;
    const float il = 1.0f / sqrtf(l2);
    n[0] = nx * il; n[1] = ny * il; n[2] = nz * il;
    return true;
}

// Motion of the intersection line of a moving plane (normal n_var) with a
// fixed one (n_fix), projected to a 2D slice (src/cuda/rasterize.cu:636-651).
__device__ __forceinline__ void slice_motion(float nvx, float nvy, float nfx, float nfy, float& dbx,
                                             float& dby) {
    float il = 1.0f / sqrtf(fmaxf(1e-30f, nvx * nvx + nvy * nvy));
    nvx *= il; nvy *= il;
    il = 1.0f / sqrtf(fmaxf(1e-30f, nfx * nfx + nfy * nfy));
    nfx *= il; nfy *= il;
    const float bx = -nfy, by = nfx;
    const float den = bx * nvx + by * nvy;
    const float s = (fabsf(den) > 1e-12f) ? (bx / den) : 0.f;
    dbx = s * nvx;
    dby = s * nvy;
}

// Spread an NDC-space motion (gx, gy, gz) at a point with barycentrics
// (b0, b1) onto the triangle's clip-space corners.
__device__ __forceinline__ void spread(float g[12], float b0, float b1, float clip_z, float clip_w,
                                       float mid_x, float mid_y, float gx, float gy, float gz, bool with_z) {
    const float b2 = 1.f - b0 - b1;
    const float cx = mid_x * clip_w, cy = mid_y * clip_w;
    const float iw = 1.f / clip_w;
    const float iw2 = iw * iw;
    const float gs = with_z ? gx * cx + gy * cy + gz * clip_z : gx * cx + gy * cy;
    const float b[3] = {b0, b1, b2};
    for (int k = 0; k < 3; ++k) {
        g[k * 4 + 0] += b[k] * gx * iw;
        g[k * 4 + 1] += b[k] * gy * iw;
        if (with_z) g[k * 4 + 2] += b[k] * gz * iw;
        g[k * 4 + 3] += -b[k] * gs * iw2;
    }
}

__device__ __forceinline__ bool vertices_ok(const int vi[3], int V) {
    return vi[0] >= 0 && vi[0] < V && vi[1] >= 0 && vi[1] < V && vi[2] >= 0 && vi[2] < V;
}

// Per-pixel edge gradient.  With REC the triangle corners come from the
// raster's per-(view, triangle) clip records (one 48-byte read instead of an
// index read followed by three vertex reads); b >= 0 gives the pixel's view.
// With KNOWN the caller also gives the visible triangle it0 (a covered
// pixel) and C is 3: every load then depends on (p0, b, it0) only - the five
// raster texels, the record and the five pixels' colours and gradients are
// all in flight together.
template <bool REC, bool KNOWN = false>
__device__ __forceinline__ bool edge_eval(const Views& v, int T, int C, const float* __restrict__ color,
                                          const float* __restrict__ grad_color, const float4* __restrict__ rast,
                                          int V, const float4* __restrict__ pos, const int* __restrict__ tri,
                                          const TriRecord* __restrict__ recs, long long p0, int& b, float g[12],
                                          int v0[3], int& it0) {
    const float4 r0 = rast[p0];
    if constexpr (!KNOWN) {
        it0 = static_cast<int>(r0.w) - 1;
        if (it0 < 0 || it0 >= T) return false;
        if (b < 0) {
            b = 0;
            while (b + 1 < v.n && v.pix_base[b + 1] <= p0) ++b;
        }
    }
    const int W = v.W[b], H = v.H[b];
    const int local = static_cast<int>(p0 - v.pix_base[b]);
    const int y0 = local / W, x0 = local - y0 * W;
    const float nx0 = pixel_ndc(x0, W), ny0 = pixel_ndc(y0, H);
    const float depth0 = r0.z;

    const int DX[4] = {1, -1, 0, 0};
    const int DY[4] = {0, 0, -1, 1};
    // neighbour raster texels first: independent loads in flight together
    bool inb[4];
    long long pn[4];
    float4 rn[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int x1 = x0 + DX[e], y1 = y0 + DY[e];
        inb[e] = !(y1 < 0 || y1 >= H || x1 < 0 || x1 >= W);
        pn[e] = inb[e] ? v.pix_base[b] + (long long)y1 * W + x1 : p0;
        rn[e] = rast[pn[e]];
    }

    const float4* bp = pos + (long long)b * V;
    float4 P0[3];
    if constexpr (REC) {
        const TriRecord& rc = recs[(long long)b * T + it0];
        P0[0] = rc.p0; P0[1] = rc.p1; P0[2] = rc.p2;
        v0[0] = tri[3 * it0]; v0[1] = tri[3 * it0 + 1]; v0[2] = tri[3 * it0 + 2];
    } else {
        v0[0] = tri[3 * it0]; v0[1] = tri[3 * it0 + 1]; v0[2] = tri[3 * it0 + 2];
        if (!vertices_ok(v0, V)) return false;
        P0[0] = bp[v0[0]]; P0[1] = bp[v0[1]]; P0[2] = bp[v0[2]];
    }
    const float* c0 = color + p0 * C;
    const float* g0 = grad_color + p0 * C;
    // KNOWN: the colour adjoints of the four edges from preloaded values
    // (the same sums, in the same order, as the generic loop below)
    float adjv[4] = {0.f, 0.f, 0.f, 0.f};
    if constexpr (KNOWN) {
        float cq[5][3], gq[5][3];
#pragma unroll
        for (int e = 0; e < 5; ++e) {
            const long long q = e == 0 ? p0 : pn[e - 1];
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                cq[e][i] = color[q * 3 + i];
                gq[e][i] = grad_color[q * 3 + i];
            }
        }
#pragma unroll
        for (int e = 0; e < 4; ++e)
#pragma unroll
            for (int i = 0; i < 3; ++i) adjv[e] += 0.5f * (gq[0][i] + gq[e + 1][i]) * (cq[0][i] - cq[e + 1][i]);
    }

    for (int k = 0; k < 12; ++k) g[k] = 0.f;
    bool touched = false;

#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int x1 = x0 + DX[e], y1 = y0 + DY[e];
        if (!inb[e]) continue;
        const long long p1 = pn[e];
        const float4 r1 = rn[e];
        const float depth1 = r1.z;
        int it1 = static_cast<int>(r1.w) - 1;
        if (it1 >= T) it1 = -1;
        float adj = 0.f;
        if constexpr (KNOWN) {
            adj = adjv[e];
        } else {
            const float* c1 = color + p1 * C;
            const float* g1 = grad_color + p1 * C;
            for (int i = 0; i < C; ++i) adj += 0.5f * (g0[i] + g1[i]) * (c0[i] - c1[i]);
        }
        if (adj == 0.f) continue;

        const float gx = adj * static_cast<float>(DX[e]) * static_cast<float>(W) / 2.f;
        const float gy = adj * static_cast<float>(DY[e]) * static_cast<float>(H) / 2.f;
        const float nx1 = pixel_ndc(x1, W), ny1 = pixel_ndc(y1, H);
        const float mx = 0.5f * (nx0 + nx1), my = 0.5f * (ny0 + ny1);

        Hit h;
        if (it1 == it0) {
            // interior: the boundary does not change triangles.  The
            // neighbour, covered by the same triangle, evaluates this pair
            // too and gets the same contribution bit for bit (the adjoint
            // and the step direction both change sign; the midpoint is
            // shared), so the -x / -y pixel of the pair skips it and the
            // +x / +y pixel adds it twice (exactly 2x, by linearity).
            if (DX[e] + DY[e] < 0) continue;
            if (cover(my, mx, P0[0], P0[1], P0[2], h)) {
                spread(g, h.b0, h.b1, h.z, h.w, mx, my, 2.f * gx, 2.f * gy, 0.f, false);
                touched = true;
            }
            continue;
        }
        Hit h1;
        const bool own_covers_1 = cover(ny1, nx1, P0[0], P0[1], P0[2], h1);
        if (own_covers_1) {
            // implicit edge: triangle intersection or opacity edge
            const float depth1_own = h1.z / h1.w;
            if (it1 >= 0 && depth1 < depth1_own) {
                float4 P1[3];
                if constexpr (REC) {
                    const TriRecord& rc = recs[(long long)b * T + it1];
                    P1[0] = rc.p0; P1[1] = rc.p1; P1[2] = rc.p2;
                } else {
                    const int v1[3] = {tri[3 * it1], tri[3 * it1 + 1], tri[3 * it1 + 2]};
                    if (!vertices_ok(v1, V)) continue;
                    P1[0] = bp[v1[0]]; P1[1] = bp[v1[1]]; P1[2] = bp[v1[2]];
                }
                Hit h0;
                if (cover(ny0, nx0, P1[0], P1[1], P1[2], h0)) {
                    const float depth0_other = h0.z / h0.w;
                    if (depth0 < depth0_other) {
                        float nvar[3], nfix[3];
                        if (ndc_normal(P0, nvar) && ndc_normal(P1, nfix)) {
                            float mgx = 0.f, mgy = 0.f, mgz = 0.f;
                            if (DX[e] != 0) {
                                float dbx, dbz;
                                slice_motion(nvar[0], nvar[2], nfix[0], nfix[2], dbx, dbz);
                                mgx = gx * dbx;
                                mgz = gx * dbz;
                            } else {
                                float dby, dbz;
                                slice_motion(nvar[1], nvar[2], nfix[1], nfix[2], dby, dbz);
                                mgy = gy * dby;
                                mgz = gy * dbz;
                            }
                            if (cover(my, mx, P0[0], P0[1], P0[2], h)) {
                                spread(g, h.b0, h.b1, h.z, h.w, mx, my, mgx, mgy, mgz, true);
                                touched = true;
                            }
                        }
                    }
                }
            } else if (it1 < 0 || depth0 < depth1) {
                // active opacity edge
                if (cover(my, mx, P0[0], P0[1], P0[2], h)) {
                    spread(g, h.b0, h.b1, h.z, h.w, mx, my, gx, gy, 0.f, false);
                    touched = true;
                }
            }
        } else if (it1 < 0 || depth0 < depth1) {
            // active geometric edge of the pixel's own triangle
            int i0, i1;
            float s0, s1;
            if (crossed_edge(P0, nx0, ny0, nx1, ny1, i0, i1, s0, s1)) {
                // the crossed edge's end points, selected without dynamic
                // register indexing (which would live in scratch memory)
                const float4 A = i0 == 0 ? P0[0] : (i0 == 1 ? P0[1] : P0[2]);
                const float4 Bv = i1 == 0 ? P0[0] : (i1 == 1 ? P0[1] : P0[2]);
                const float cx = s0 * A.x + s1 * Bv.x;
                const float cy = s0 * A.y + s1 * Bv.y;
                const float cw = s0 * A.w + s1 * Bv.w;
                const float iw = 1.f / cw;
                const float iw2 = iw * iw;
                const float gs = gx * cx + gy * cy;
                const float ax = s0 * gx * iw, bx = s1 * gx * iw;
                const float ay = s0 * gy * iw, by = s1 * gy * iw;
                const float aw = -s0 * gs * iw2, bw = -s1 * gs * iw2;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    if (k == i0) { g[k * 4 + 0] += ax; g[k * 4 + 1] += ay; g[k * 4 + 3] += aw; }
                    if (k == i1) { g[k * 4 + 0] += bx; g[k * 4 + 1] += by; g[k * 4 + 3] += bw; }
                }
                touched = true;
            }
        }
    }
    return touched;
}

// Per-lane evaluation followed by per-lane atomics (generic entry points).
template <bool REC>
__device__ __forceinline__ void edge_at(const Views& v, int T, int C, const float* __restrict__ color,
                                        const float* __restrict__ grad_color, const float4* __restrict__ rast,
                                        int V, const float4* __restrict__ pos, float* __restrict__ grad_pos,
                                        const int* __restrict__ tri, const TriRecord* __restrict__ recs,
                                        long long p0, int b) {
    float g[12];
    int v0[3] = {0, 0, 0};
    int it0 = -1;
    if (!edge_eval<REC>(v, T, C, color, grad_color, rast, V, pos, tri, recs, p0, b, g, v0, it0)) return;
    float* gp = grad_pos + (long long)b * V * 4;
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (g[k * 4 + j] != 0.f) unsafeAtomicAdd(&gp[v0[k] * 4 + j], g[k * 4 + j]);
}

__global__ void k_edge_grad(Views v, int T, int C, const float* __restrict__ color,
                            const float* __restrict__ grad_color, const float4* __restrict__ rast,
                            int V, const float4* __restrict__ pos, float* __restrict__ grad_pos,
                            const int* __restrict__ tri) {
    const long long p0 = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (p0 >= v.pix_base[v.n]) return;
    edge_at<false>(v, T, C, color, grad_color, rast, V, pos, grad_pos, tri, nullptr, p0, -1);
}

// Edge gradients over the covered pixels listed in tile order (each screen
// tile's covered pixels together, so a tile's raster, colour and gradient
// rows are fetched once and its neighbours' rows are near), entries
// (pixel, view, visible triangle, 0); the count is read on the device.  Consecutive entries mostly share a few triangles of
// one view, so a wave groups its lanes by (view, gradient triangle), adds
// each lane's 12 clip-space components into its group's LDS slot, and
// flushes one global atomic per (group, component).  A persistent grid
// strides over the list, so every lane of a wave has a pixel except in the
// last stride.
__global__ void __launch_bounds__(256)
k_edge_grad_sorted(Views v, int T, int C, const float* __restrict__ color, const float* __restrict__ grad_color,
                   const float4* __restrict__ rast, int V, const float4* __restrict__ pos,
                   float* __restrict__ grad_pos, const int* __restrict__ tri, const TriRecord* __restrict__ recs,
                   const int4* __restrict__ px_list, const int* __restrict__ n_list) {
    __shared__ float slot[4][64][13];
    __shared__ long long sv[4][64][3];
    const int n = *n_list;
    const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
    const long long stride = (long long)gridDim.x * 256;
    // the loop bound depends on the wave's first index only: wave-uniform
    for (long long base = (long long)blockIdx.x * 256 + w * 64; base < n; base += stride) {
        const long long i = base + lane;
        float g[12];
        int v0[3] = {0, 0, 0};
        int it0 = -1, b = -1;
        bool act = false;
        if (i < n) {
            const int4 e = px_list[i];
            b = e.y;
            it0 = e.z;
            act = edge_eval<true, true>(v, T, 3, color, grad_color, rast, V, pos, tri, recs, e.x, b, g, v0, it0);
        }
        const long long key = act ? (long long)b * T + it0 : -1;
        int grp = -1, ng = 0;
        unsigned long long todo = __ballot(act);
        while (todo) {
            const int leader = __ffsll(static_cast<long long>(todo)) - 1;
            // the leader is wave-uniform: its key by two scalar lane reads
            const long long k = static_cast<long long>(
                static_cast<unsigned long long>(static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(key), leader))) |
                static_cast<unsigned long long>(static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(key >> 32), leader))) << 32);
            const unsigned long long same = __ballot(act && key == k);
            if (act && key == k) grp = ng;
            if (lane == leader)
                for (int c = 0; c < 3; ++c) sv[w][ng][c] = (long long)b * V + v0[c];
            ++ng;
            todo &= ~same;
        }
        for (int q = lane; q < ng * 13; q += 64) (&slot[w][0][0])[q] = 0.f;
        wave_lds_sync();
        if (act)
#pragma unroll
            for (int k = 0; k < 12; ++k)
                if (g[k] != 0.f) atomicAdd(&slot[w][grp][k], g[k]);
        wave_lds_sync();
        for (int q = lane; q < ng * 12; q += 64) {
            const int gq = q / 12, k = q - gq * 12;
            const float x = slot[w][gq][k];
            if (x != 0.f) unsafeAtomicAdd(&grad_pos[sv[w][gq][k / 4] * 4 + (k % 4)], x);
        }
        wave_lds_sync();   // the slots are reused by the next stride
    }
}

}  // namespace

hipError_t launch_edge_grad_sorted(const Views& v, int C, const float* color, const float* grad_color,
                                   const float* rast, int V, const float* pos, float* grad_pos, int T,
                                   const int* tri, const TriRecord* recs, const int4* px_list, const int* n_list,
                                   long long max_n, hipStream_t s) {
    if (max_n <= 0) return hipSuccess;
    if (C != 3) return hipErrorInvalidValue;
    const long long blocks = std::min<long long>((max_n + 255) / 256, 2048);
    k_edge_grad_sorted<<<static_cast<unsigned>(blocks), 256, 0, s>>>(
        v, T, C, color, grad_color, reinterpret_cast<const float4*>(rast), V, reinterpret_cast<const float4*>(pos),
        grad_pos, tri, recs, px_list, n_list);
    return hipGetLastError();
}

hipError_t launch_edge_grad(const Views& v, int C, const float* color, const float* grad_color,
                            const float* rast, int V, const float* pos, float* grad_pos, int T,
                            const int* tri, hipStream_t s) {
    const long long P = v.pix_base[v.n];
    if (P <= 0) return hipSuccess;
    k_edge_grad<<<static_cast<unsigned>((P + 255) / 256), 256, 0, s>>>(
        v, T, C, color, grad_color, reinterpret_cast<const float4*>(rast), V,
        reinterpret_cast<const float4*>(pos), grad_pos, tri);
    return hipGetLastError();
}

}  // namespace ds
