// Per-triangle visible-pixel counts of a raster: the bincount of the 1-based
// triangle ids that count_triangle_ids computes with torch
// (python/diffsoup/rasterize.py:348-366), used by the visibility pruning of
// examples/utils.py:445-490.  Adjacent pixels mostly share a triangle, so a
// wave groups its lanes by id and issues one 64-bit atomic per distinct id.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "diffsoup_hip.h"

namespace ds {
namespace {

constexpr int kWave = 64;

__global__ void k_zero_i64(long long* p, long long n) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = 0;
}

__global__ void k_count_ids(long long P, const float4* __restrict__ rast, int T, long long* __restrict__ counts) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    const int lane = threadIdx.x % kWave;
    // the loop bound is uniform over a wave (it depends on the wave's base
    // index only), so every lane takes part in the ballots below
    const long long base0 = (blockIdx.x * (long long)blockDim.x + threadIdx.x) - lane;
    for (long long base = base0; base < P; base += stride) {
        const long long p = base + lane;
        int id = -1;
        if (p < P) {
            const int t = static_cast<int>(rast[p].w) - 1;
            if (t >= 0 && t < T) id = t;
        }
        unsigned long long todo = __ballot(id >= 0);
        while (todo) {
            const int leader = __ffsll(static_cast<long long>(todo)) - 1;
            const int lid = __shfl(id, leader, kWave);
            const unsigned long long same = __ballot(id == lid);
            if (lane == leader) atomicAdd(reinterpret_cast<unsigned long long*>(counts + lid),
                                          static_cast<unsigned long long>(__popcll(same)));
            todo &= ~same;
        }
    }
}

}  // namespace

int visibility_fail(int code, const char* msg);

}  // namespace ds

extern "C" int ds_count_triangle_ids(int64_t P, const float* rast, int T, int64_t* counts, int accumulate,
                                     void* stream) {
    using namespace ds;
    if (P < 0 || T < 0 || (P > 0 && !rast) || (T > 0 && !counts))
        return visibility_fail(DS_ERR_INVALID, "count_triangle_ids: bad sizes or null pointer");
    hipStream_t s = static_cast<hipStream_t>(stream);
    long long* c = reinterpret_cast<long long*>(counts);
    if (!accumulate && T > 0)
        k_zero_i64<<<static_cast<unsigned>(std::min<long long>((T + 255) / 256, 4096)), 256, 0, s>>>(c, T);
    if (P > 0 && T > 0) {
        const long long blocks = std::min<long long>((P + 255) / 256, 8192);
        k_count_ids<<<static_cast<unsigned>(blocks), 256, 0, s>>>(P, reinterpret_cast<const float4*>(rast), T, c);
    }
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? DS_OK : visibility_fail(DS_ERR_HIP, hipGetErrorString(e));
}
