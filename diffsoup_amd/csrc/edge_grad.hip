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
    if (segments_meet(x0, y0, x1, y1, rx[0], ry[0], rx[1], ry[1], ix, iy)) { i0 = 0; i1 = 1; }
    else if (segments_meet(x0, y0, x1, y1, rx[1], ry[1], rx[2], ry[2], ix, iy)) { i0 = 1; i1 = 2; }
    else if (segments_meet(x0, y0, x1, y1, rx[2], ry[2], rx[0], ry[0], ix, iy)) { i0 = 2; i1 = 0; }
    else return false;
    const float4 A = P[i0], B = P[i1];
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

__global__ void k_edge_grad(Views v, int T, int C, const float* __restrict__ color,
                            const float* __restrict__ grad_color, const float4* __restrict__ rast,
                            int V, const float4* __restrict__ pos, float* __restrict__ grad_pos,
                            const int* __restrict__ tri) {
    const long long p0 = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (p0 >= v.pix_base[v.n]) return;
    const float4 r0 = rast[p0];
    const int it0 = static_cast<int>(r0.w) - 1;
    if (it0 < 0 || it0 >= T) return;
    int b = 0;
    while (b + 1 < v.n && v.pix_base[b + 1] <= p0) ++b;
    const int W = v.W[b], H = v.H[b];
    const int local = static_cast<int>(p0 - v.pix_base[b]);
    const int y0 = local / W, x0 = local - y0 * W;
    const float nx0 = pixel_ndc(x0, W), ny0 = pixel_ndc(y0, H);
    const float depth0 = r0.z;

    const float4* bp = pos + (long long)b * V;
    const int v0[3] = {tri[3 * it0], tri[3 * it0 + 1], tri[3 * it0 + 2]};
    if (!vertices_ok(v0, V)) return;
    const float4 P0[3] = {bp[v0[0]], bp[v0[1]], bp[v0[2]]};
    const float* c0 = color + p0 * C;
    const float* g0 = grad_color + p0 * C;

    float g[12];
    for (int k = 0; k < 12; ++k) g[k] = 0.f;
    bool touched = false;

    const int DX[4] = {1, -1, 0, 0};
    const int DY[4] = {0, 0, -1, 1};
    for (int e = 0; e < 4; ++e) {
        const int x1 = x0 + DX[e], y1 = y0 + DY[e];
        if (y1 < 0 || y1 >= H || x1 < 0 || x1 >= W) continue;
        const long long p1 = v.pix_base[b] + (long long)y1 * W + x1;
        const float4 r1 = rast[p1];
        const float depth1 = r1.z;
        int it1 = static_cast<int>(r1.w) - 1;
        if (it1 >= T) it1 = -1;
        const float* c1 = color + p1 * C;
        const float* g1 = grad_color + p1 * C;
        float adj = 0.f;
        for (int i = 0; i < C; ++i) adj += 0.5f * (g0[i] + g1[i]) * (c0[i] - c1[i]);
        if (adj == 0.f) continue;

        const float gx = adj * static_cast<float>(DX[e]) * static_cast<float>(W) / 2.f;
        const float gy = adj * static_cast<float>(DY[e]) * static_cast<float>(H) / 2.f;
        const float nx1 = pixel_ndc(x1, W), ny1 = pixel_ndc(y1, H);
        const float mx = 0.5f * (nx0 + nx1), my = 0.5f * (ny0 + ny1);

        Hit h;
        if (it1 == it0) {
            // interior: the boundary does not change triangles
            if (cover(my, mx, P0[0], P0[1], P0[2], h)) {
                spread(g, h.b0, h.b1, h.z, h.w, mx, my, gx, gy, 0.f, false);
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
                const int v1[3] = {tri[3 * it1], tri[3 * it1 + 1], tri[3 * it1 + 2]};
                if (!vertices_ok(v1, V)) continue;
                const float4 P1[3] = {bp[v1[0]], bp[v1[1]], bp[v1[2]]};
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
                const float cx = s0 * P0[i0].x + s1 * P0[i1].x;
                const float cy = s0 * P0[i0].y + s1 * P0[i1].y;
                const float cw = s0 * P0[i0].w + s1 * P0[i1].w;
                const float iw = 1.f / cw;
                const float iw2 = iw * iw;
                const float gs = gx * cx + gy * cy;
                g[i0 * 4 + 0] += s0 * gx * iw;
                g[i1 * 4 + 0] += s1 * gx * iw;
                g[i0 * 4 + 1] += s0 * gy * iw;
                g[i1 * 4 + 1] += s1 * gy * iw;
                g[i0 * 4 + 3] += -s0 * gs * iw2;
                g[i1 * 4 + 3] += -s1 * gs * iw2;
                touched = true;
            }
        }
    }
    if (!touched) return;
    float* gp = grad_pos + (long long)b * V * 4;
    for (int k = 0; k < 3; ++k)
        for (int j = 0; j < 4; ++j)
            if (g[k * 4 + j] != 0.f) unsafeAtomicAdd(&gp[v0[k] * 4 + j], g[k * 4 + j]);
}

}  // namespace

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
