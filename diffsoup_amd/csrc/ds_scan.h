// Prefix sums of int32 counts (scan.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

namespace ds {

size_t scan_tmp_bytes(long long n);
// Device slot of tmp that holds the total after any scan of n items.
long long* scan_total_slot(void* tmp, long long n);
// out[0..n]: exclusive prefix, out[n] = total.
hipError_t launch_exclusive_scan(const int* in, long long n, int* out, void* tmp, hipStream_t s);
// In-place exclusive prefix of data[0..n).
hipError_t launch_exclusive_scan_inplace(int* data, long long n, void* tmp, hipStream_t s);
// out[i] = sum of in[0..i] as int64.
hipError_t launch_inclusive_scan_i64(const int* in, long long n, long long* out, void* tmp, hipStream_t s);

}  // namespace ds
