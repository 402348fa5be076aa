// Library runtime: error state, strided-geometry helper, twiddle-table cache.
#include <math.h>
#include <stdarg.h>

#include <map>
#include <mutex>
#include <utility>
#include <vector>

#include "common.h"

namespace sfa {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }

int fail(int code, const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int make_geom(int ndim, const int64_t* shape, const int64_t* strides, int axis, Geom* g)
{
    SFA_CHECK(ndim >= 1 && ndim <= 3, SFA_ERR_ARG, "ndim must be 1..3 (got %d)", ndim);
    SFA_CHECK(axis >= 0 && axis < ndim, SFA_ERR_ARG, "axis %d out of range for ndim %d", axis, ndim);
    for (int i = 0; i < ndim; ++i)
        SFA_CHECK(shape[i] >= 1, SFA_ERR_ARG, "shape[%d] = %lld must be >= 1", i, (long long)shape[i]);
    g->n = shape[axis];
    g->sa = strides[axis];
    int64_t od[2] = {1, 1}, os[2] = {0, 0};
    int j = 0;
    for (int i = 0; i < ndim; ++i) {
        if (i == axis) continue;
        od[j] = shape[i];
        os[j] = strides[i];
        ++j;
    }
    // Line order is C order over the non-axis dims: (d1 slow, d2 fast).
    g->d1 = od[0];
    g->s1 = os[0];
    g->d2 = od[1];
    g->s2 = os[1];
    if (j == 1) {  // 2D: one other dim -> it is the fast one
        g->d1 = 1;
        g->s1 = 0;
        g->d2 = od[0];
        g->s2 = os[0];
    } else if (j == 0) {
        g->d1 = 1;
        g->s1 = 0;
        g->d2 = 1;
        g->s2 = 0;
    }
    return SFA_OK;
}

// ------------------------------------------------------------- allocation guard
static thread_local bool g_no_alloc = false;

bool alloc_forbidden() { return g_no_alloc; }
NoAllocScope::NoAllocScope() : prev(g_no_alloc) { g_no_alloc = true; }
NoAllocScope::~NoAllocScope() { g_no_alloc = prev; }

// ------------------------------------------------------------- twiddles
static std::mutex g_tab_mu;
static std::map<std::pair<int, int>, std::pair<int, double2*>> g_tabs;  // (kind, n) -> (device, ptr)

static const double2* get_table(int kind, int n)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lk(g_tab_mu);
    auto key = std::make_pair(kind * 64 + dev, n);
    auto it = g_tabs.find(key);
    if (it != g_tabs.end()) return it->second.second;
    if (g_no_alloc) {
        fail(SFA_ERR_ARG, "twiddle table of length %d was not built before a no-allocation call", n);
        return nullptr;
    }
    std::vector<double2> h(n);
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int m = 0; m < n; ++m) {
        long double ang = (kind == 0) ? (-2.0L * pi * m / n) : (-pi * m / (2.0L * n));
        h[m] = make_double2((double)cosl(ang), (double)sinl(ang));
    }
    double2* d = nullptr;
    if (hipMalloc(&d, sizeof(double2) * (size_t)n) != hipSuccess) return nullptr;
    if (hipMemcpy(d, h.data(), sizeof(double2) * (size_t)n, hipMemcpyHostToDevice) != hipSuccess) {
        hipFree(d);
        return nullptr;
    }
    g_tabs[key] = std::make_pair(dev, d);
    return d;
}

const double2* fft_table(int n) { return get_table(0, n); }
const double2* quarter_table(int N) { return get_table(1, N); }

int device_cu_count()
{
    static int cached[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    if (cached[dev] == 0) {
        int cu = 0;
        if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cu <= 0)
            cu = 256;
        cached[dev] = cu;
    }
    return cached[dev];
}

}  // namespace sfa

extern "C" int sfa_version(void) { return 100; }
extern "C" const char* sfa_last_error(void) { return sfa::g_err.c_str(); }
