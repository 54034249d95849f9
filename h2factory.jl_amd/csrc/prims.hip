// Device-wide sort and scan for the plan builder (plan.hip), on rocPRIM.  Each call allocates its
// temporary storage and returns once the result is on the device (stream-ordered, then synced).
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "h2g_internal.h"

namespace h2g {

namespace {
template <class F>
void with_temp(F &&call, hipStream_t s) {
    size_t bytes = 0;
    H2G_HIP(call(nullptr, bytes));
    DevBuf tmp;
    tmp.alloc(bytes ? bytes : 16);
    H2G_HIP(call(tmp.ptr, bytes));
    H2G_HIP(hipStreamSynchronize(s));
}
}  // namespace

void prim_sort_keys_u64(const uint64_t *in, uint64_t *out, int64_t n, int end_bit, hipStream_t s) {
    if (n <= 0) return;
    with_temp([&](void *t, size_t &b) {
        return rocprim::radix_sort_keys(t, b, in, out, (size_t)n, 0u, (unsigned)end_bit, s);
    }, s);
}

void prim_sort_pairs_u64(const uint64_t *kin, uint64_t *kout, const int64_t *vin, int64_t *vout, int64_t n,
                         int end_bit, hipStream_t s) {
    if (n <= 0) return;
    with_temp([&](void *t, size_t &b) {
        return rocprim::radix_sort_pairs(t, b, kin, kout, vin, vout, (size_t)n, 0u, (unsigned)end_bit, s);
    }, s);
}

void prim_inclusive_scan_i64(const int64_t *in, int64_t *out, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    with_temp([&](void *t, size_t &b) {
        return rocprim::inclusive_scan(t, b, in, out, (size_t)n, rocprim::plus<int64_t>(), s);
    }, s);
}

void prim_exclusive_scan_i64(const int64_t *in, int64_t *out, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    with_temp([&](void *t, size_t &b) {
        return rocprim::exclusive_scan(t, b, in, out, (int64_t)0, (size_t)n, rocprim::plus<int64_t>(), s);
    }, s);
}

}  // namespace h2g
