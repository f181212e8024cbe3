// Error reporting and device queries shared by every entry point.
#include <algorithm>
#include <mutex>
#include <vector>

#include "fm_common.h"

namespace fm {

static thread_local std::string g_last_error;

void set_error(const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

int64_t balanced_splits(int64_t N, int64_t chunk, int64_t base, int64_t s_min, int64_t capacity, int64_t s_max)
{
    const int64_t n = std::max<int64_t>(N, 1);
    int64_t s_top = cdiv(n, chunk);  // one chunk per split at most
    if (s_max > 0) s_top = std::max<int64_t>(1, std::min(s_top, s_max));
    s_min = std::max<int64_t>(1, std::min(s_min, s_top));
    capacity = std::max<int64_t>(capacity, 1);
    int64_t best = s_min, best_cost = -1;
    for (int64_t s = s_min; s <= std::min(2 * s_min, s_top); ++s) {
        const int64_t nps = cdiv(cdiv(n, s), chunk) * chunk;
        const int64_t cost = cdiv(base * cdiv(n, nps), capacity) * nps;
        if (best_cost < 0 || cost < best_cost) {
            best = s;
            best_cost = cost;
        }
    }
    return best;
}

int device_cu_count()
{
    static std::mutex mu;
    static std::vector<int> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    std::lock_guard<std::mutex> lock(mu);
    if ((int)cache.size() <= dev) cache.resize(dev + 1, 0);
    if (cache[dev] == 0) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        cache[dev] = cus;
    }
    return cache[dev];
}

__global__ void k_zero_f64(int64_t n, double *__restrict__ x)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0.0;
}

int zero_f64_async(double *x, int64_t n, hipStream_t st)
{
    if (n <= 0) return FM_OK;
    k_zero_f64<<<(unsigned)cdiv(n, 256), 256, 0, st>>>(n, x);
    FM_LAUNCH_CHECK("k_zero_f64");
    return FM_OK;
}

}  // namespace fm

extern "C" const char *fm_last_error(void) { return fm::g_last_error.c_str(); }

extern "C" int fm_version(void) { return 1; }
