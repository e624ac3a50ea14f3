// Per-pixel view direction -> real spherical-harmonic basis up to degree 2.
// Reference: src/cuda/encoding.cu:9-130.  The direction points from the far
// plane point to the near plane point of the pixel's ray (toward the camera),
// unprojected with a row-major inverse MVP.  Normalisation uses a correctly
// rounded 1/sqrt instead of an approximate reciprocal square root.
#include "ds_kernels.h"
#include "ds_math.h"

namespace ds {
namespace {

__global__ void k_sh2(Views v, const float4* __restrict__ rast, const float* __restrict__ inv_mvp,
                      float* __restrict__ enc) {
    const long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (p >= v.pix_base[v.n]) return;
    int b = 0;
    while (b + 1 < v.n && v.pix_base[b + 1] <= p) ++b;
    const int local = static_cast<int>(p - v.pix_base[b]);
    const int W = v.W[b], H = v.H[b];
    const int y = local / W, x = local - y * W;
    float* e = enc + p * 9;
    if (static_cast<int>(rast[p].w) - 1 < 0) {
        for (int i = 0; i < 9; ++i) e[i] = 0.f;
        return;
    }
    const float nx = pixel_ndc(x, W), ny = pixel_ndc(y, H);
    float sh[9];
    sh2_basis(inv_mvp + 16 * b, nx, ny, sh);
    for (int i = 0; i < 9; ++i) e[i] = sh[i];
}

}  // namespace

hipError_t launch_sh2(const Views& v, const float* rast, const float* inv_mvp, float* enc,
                      hipStream_t s) {
    const long long P = v.pix_base[v.n];
    if (P <= 0) return hipSuccess;
    k_sh2<<<static_cast<unsigned>((P + 255) / 256), 256, 0, s>>>(
        v, reinterpret_cast<const float4*>(rast), inv_mvp, enc);
    return hipGetLastError();
}

}  // namespace ds
