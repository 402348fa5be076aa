// Fused (alpha - lap) u = f solve on Chebyshev(SD/SN) x C2C x R2C, FP64.
//
// Pipeline over the complex workspace W (Nx, Ny, Nz/2+1):
//   z: R2C (scale 1/Nz)            f -> W
//   y: C2C forward (scale 1/Ny)    W -> W
//   x: fused x-pass                W -> W
//        DCT-II * pi/(2Nx) -> composite stencil -> Helmholtz LU computed on
//        the fly (alfa = -1, beta = alpha + ky^2 + kz^2) + solve ->
//        to_ortho -> DCT-III * 0.5 + t_0/2, all in LDS
//   y: C2C inverse                 W -> W
//   z: C2R                         W -> u
#include <math.h>

#include <string.h>

#include <algorithm>
#include <chrono>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "fft_device.h"
#include "helm_scan.h"

int fft_c2r_add(const void* in, double* out, int ndim, const int64_t* real_shape, const int64_t* in_strides,
                const int64_t* out_strides, int axis, double scale, void* stream, const double* add,
                int64_t add_rows);

namespace sfa {

struct HelmTables {
    const double *A0, *A2, *A4, *Bm2, *B0, *B2;  // unscaled diagonals, M = N-2
};

// Closed-form Chebyshev SD/SN Galerkin diagonals (chebyshev/matrices.py:221-281,
// 377-427, 1047-1165), uploaded once per (N, neumann) and kept for the process.
static std::mutex g_helm_mu;
static std::map<std::pair<int, int>, double*> g_helm_tabs;  // (N*2+neumann, device) -> 6*M doubles

static const double* helm_tables(int N, int neumann)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lk(g_helm_mu);
    auto key = std::make_pair(N * 2 + neumann, dev);
    auto it = g_helm_tabs.find(key);
    if (it != g_helm_tabs.end()) return it->second;
    if (alloc_forbidden()) return nullptr;
    const int M = N - 2;
    const double pi = 3.14159265358979323846;
    std::vector<double> h(6 * (size_t)M, 0.0);
    double* A0 = h.data();
    double* A2 = A0 + M;
    double* A4 = A2 + M;
    double* Bm2 = A4 + M;
    double* B0 = Bm2 + M;
    double* B2 = B0 + M;
    std::vector<double> hn(N, pi / 2);  // Chebyshev GC norms
    hn[0] = pi;
    for (int k = 0; k < M; ++k) {
        if (!neumann) {
            A0[k] = -2 * pi * (k + 1) * (k + 2);
            if (k < M - 2) A2[k] = -4 * pi * (k + 1);
            if (k < M - 4) A4[k] = -4 * pi * (k + 1);
            B0[k] = hn[k] + hn[k + 2];
            if (k < M - 2) { B2[k] = -hn[1]; Bm2[k] = -hn[1]; }
        } else {
            const double kk = k;
            A0[k] = -2 * pi * kk * kk * (kk + 1) / (kk + 2);
            if (k < M - 2) A2[k] = -4 * pi * (kk + 2) * (kk + 2) * (kk + 1) / ((kk + 2) * (kk + 2));
            if (k < M - 4) A4[k] = -4 * pi * (kk + 4) * (kk + 4) * (kk + 1) / ((kk + 2) * (kk + 2));
            const double al = (kk / (kk + 2)) * (kk / (kk + 2));
            const double al2 = ((kk + 2) / (kk + 4)) * ((kk + 2) / (kk + 4));
            (void)al2;
            B0[k] = hn[k] + al * al * hn[k + 2];
            if (k < M - 2) { B2[k] = -al * hn[k + 2]; Bm2[k] = B2[k]; }
        }
    }
    double* d = nullptr;
    if (hipMalloc(&d, sizeof(double) * h.size()) != hipSuccess) return nullptr;
    if (hipMemcpy(d, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice) != hipSuccess) {
        hipFree(d);
        return nullptr;
    }
    g_helm_tabs[key] = d;
    return d;
}

struct XPassArgs {
    double2* W;
    int64_t sa, s1;  // x stride, y stride (z stride 1)
    int64_t d1, d2;  // Ny, Nz/2+1
    int N, ldn, T;
    int neumann;
    double alpha;
    int Ny;
    int E;                 // rows per lane of a 16-lane chain
    const double* K2;      // N-2 stencil entries
    const double* tabs;    // 6*(N-2) diagonals
    const double2* tw;     // FFT table N
    const double2* tq;     // quarter table N
    FFTPlan plan;
};

__device__ inline int perm_dct(int i, int N) { return (i & 1) ? (N - 1 - (i >> 1)) : (i >> 1); }

__global__ void __launch_bounds__(256) xpass_kernel(XPassArgs a)
{
    extern __shared__ double2 smem[];
    const int tid = threadIdx.x, nth = blockDim.x;
    const int64_t ntile = (a.d2 + a.T - 1) / a.T;
    const int64_t jy = blockIdx.x / ntile;
    const int64_t jz0 = (blockIdx.x - jy * ntile) * a.T;
    const int Tv = (int)((a.d2 - jz0) < a.T ? (a.d2 - jz0) : a.T);
    const int N = a.N, ldn = a.ldn, M = N - 2;
    double2* A = smem;
    double2* B = smem + (size_t)a.T * ldn;
    double* t0s = reinterpret_cast<double*>(smem + (size_t)2 * a.T * ldn);
    double2* W = a.W + jy * a.s1 + jz0;

    // 1. load with the DCT-II even/odd permutation
    for (int e = tid; e < Tv * N; e += nth) {
        const int t = e % Tv, i = e / Tv;
        A[t * ldn + perm_dct(i, N)] = W[(int64_t)i * a.sa + t];
    }
    __syncthreads();
    double2* F = run_fft<false>(A, B, a.plan, ldn, Tv, a.tw, tid, nth);
    double2* G = (F == A) ? B : A;
    // 2. DCT-II post-processing and scale pi/(2N)
    const double sc = 3.14159265358979323846 / (2.0 * N);
    for (int e = tid; e < Tv * N; e += nth) {
        const int t = e / N, k = e - t * N;
        const double2 P = F[t * ldn + k];
        const double2 Q = cconj(F[t * ldn + (k == 0 ? 0 : N - k)]);
        const double2 w = a.tq[k];
        const double2 s1 = cmul(w, cadd(P, Q));
        const double2 s2 = cmul(w, csub(P, Q));
        G[t * ldn + k] = make_double2(s1.x * sc, s2.y * sc);
    }
    __syncthreads();
    // 3. composite stencil -> F
    for (int e = tid; e < Tv * N; e += nth) {
        const int t = e / N, k = e - t * N;
        double2 v = G[t * ldn + k];
        if (k < N - 2) {
            const double kk = a.K2[k];
            const double2 u = G[t * ldn + k + 2];
            v.x += kk * u.x;
            v.y += kk * u.y;
        }
        F[t * ldn + k] = v;
    }
    __syncthreads();
    // 4. Helmholtz LU on the fly + solve: 16 lanes per (line, parity) chain
    {
        const int t = tid >> 5, p = (tid >> 4) & 1, lane = tid & 15;
        if (t < Tv) {
            const int64_t jz = jz0 + t;
            const double ky = (jy < a.Ny / 2 || (a.Ny % 2 == 1 && jy == a.Ny / 2)) ? (double)jy
                                                                                 : (double)(jy - a.Ny);
            const double kz = (double)jz;
            const double be = a.alpha + ky * ky + kz * kz;
            const int M = N - 2;
            HelmTab tab{a.tabs, a.tabs + M, a.tabs + 2 * M, a.tabs + 3 * M, a.tabs + 4 * M,
                        a.tabs + 5 * M};
            double2* X = F + t * ldn;
            switch (a.E) {
                case 1: helm_chain_scan<1>(X, p, M, -1.0, be, a.neumann != 0, tab, lane); break;
                case 2: helm_chain_scan<2>(X, p, M, -1.0, be, a.neumann != 0, tab, lane); break;
                case 4: helm_chain_scan<4>(X, p, M, -1.0, be, a.neumann != 0, tab, lane); break;
                case 8: helm_chain_scan<8>(X, p, M, -1.0, be, a.neumann != 0, tab, lane); break;
                default: helm_chain_scan<16>(X, p, M, -1.0, be, a.neumann != 0, tab, lane); break;
            }
            if (p == 0 && lane == 0) {  // BC slots of a fresh solution are zero
                X[M] = make_double2(0.0, 0.0);
                X[M + 1] = make_double2(0.0, 0.0);
            }
        }
    }
    __syncthreads();
    // 5. to_ortho -> G, keep t_0
    for (int e = tid; e < Tv * N; e += nth) {
        const int t = e / N, k = e - t * N;
        double2 v = (k < N - 2) ? F[t * ldn + k] : make_double2(0.0, 0.0);
        if (k >= 2) {
            const double kk = a.K2[k - 2];
            const double2 u = F[t * ldn + k - 2];
            v.x += kk * u.x;
            v.y += kk * u.y;
        }
        G[t * ldn + k] = v;
        if (k == 0) {
            t0s[2 * t] = v.x;
            t0s[2 * t + 1] = v.y;
        }
    }
    __syncthreads();
    // 6. DCT-III pre-processing -> F
    for (int e = tid; e < Tv * N; e += nth) {
        const int t = e / N, k = e - t * N;
        const double2 ck = G[t * ldn + k];
        const double2 cn = (k == 0) ? make_double2(0.0, 0.0) : G[t * ldn + N - k];
        const double2 z = make_double2(ck.x + cn.y, ck.y - cn.x);
        F[t * ldn + k] = cmul(cconj(a.tq[k]), z);
    }
    __syncthreads();
    double2* H = run_fft<true>(F, G, a.plan, ldn, Tv, a.tw, tid, nth);
    // 7. store 0.5 y + t_0/2 with the inverse permutation
    for (int e = tid; e < Tv * N; e += nth) {
        const int t = e % Tv, i = e / Tv;
        const double2 y = H[t * ldn + perm_dct(i, N)];
        W[(int64_t)i * a.sa + t] = make_double2(0.5 * y.x + 0.5 * t0s[2 * t], 0.5 * y.y + 0.5 * t0s[2 * t + 1]);
    }
}

// K2 stencil of the Chebyshev SD/SN bases (chebyshev/bases.py:546, 937)
static std::map<std::pair<int, int>, double*> g_k2;
static const double* cheb_k2(int N, int neumann)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lk(g_helm_mu);
    auto key = std::make_pair(N * 2 + neumann, dev);
    auto it = g_k2.find(key);
    if (it != g_k2.end()) return it->second;
    if (alloc_forbidden()) return nullptr;
    std::vector<double> h(N - 2);
    for (int k = 0; k < N - 2; ++k) {
        const double kk = k;
        h[k] = neumann ? -(kk / (kk + 2)) * (kk / (kk + 2)) : -1.0;
    }
    double* d = nullptr;
    if (hipMalloc(&d, sizeof(double) * h.size()) != hipSuccess) return nullptr;
    if (hipMemcpy(d, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice) != hipSuccess) {
        hipFree(d);
        return nullptr;
    }
    g_k2[key] = d;
    return d;
}

int fast_xpass(double2* W, int Nx, int Ny, int Nzh, int Kp, double alpha, int neumann,
               const double* K2, const XPlan& pl, int64_t nb, hipStream_t s, int* rc);
int fast_xpass_mixed(void* W64, void* W32, int Nx, int Ny, int Nzh, int Kh, double alpha,
                     int neumann, const double* K2, const XPlan& pl, int64_t nb, hipStream_t s, int* rc,
                     double* stats);

// statistics delivery on its own, where the inverse y pass cannot carry it
__global__ void __launch_bounds__(64) stat_deliver_kernel(const double* stats, StatOut out)
{
    stat_deliver(stats, out);
}

int fast_c2c(const Geom& gi, const Geom& go, const void* in, void* out, int sign, double scale,
             hipStream_t s, int* rc);
int fast_c2c_f32(const Geom& gi, const Geom& go, const void* in, void* out, int sign, double scale,
                 hipStream_t s, int* rc);
int fast_c2c_mixed(const Geom& g64, void* W64, const Geom& g32, void* W32, int sign, double scale,
                   hipStream_t s, int* rc, const double* stats = nullptr,
                   StatOut out = StatOut{nullptr, nullptr, 0u});
int split_rc(bool r2c, int H, const void* in, void* o64, void* o32, int64_t nrows, int Kh,
             double scale, hipStream_t s, int* rc, double* zero = nullptr, int nzero = 0,
             const double* add = nullptr, int64_t add_rows = 1);

// ------------------------------------------------------------------ plan
// Device plan of the fused solve, the counterpart of chebyshev.la.Helmholtz's
// factor arrays (chebyshev/la.py:270-276) deduplicated over the distinct
// beta = alpha + ky^2 + kz^2: the map from a (ky, kz) line to its entry, s per
// entry and, per (entry, parity chain), the LU recurrence state at the 16
// segment starts of the chain (the solve regenerates each segment's factors
// from it).  Layout (bytes): linemap int32[nl] | s double[nu] |
// bs double4[nu][2][16].
struct PlanLayout {
    int64_t nl, nu, K, E, off_s, off_bs, bytes;
};

static int64_t round256(int64_t b) { return (b + 255) / 256 * 256; }

static void plan_shape(int64_t Nx, int64_t Ny, int64_t Nz, PlanLayout* L, std::vector<int>* map,
                       std::vector<double>* svals)
{
    const int64_t Nzh = Nz / 2 + 1;
    L->nl = Ny * Nzh;
    L->K = (Nx - 2 + 1) / 2;
    L->E = (L->K + 15) / 16;
    std::vector<int64_t> sv((size_t)L->nl);
    for (int64_t j = 0; j < Ny; ++j) {
        const int64_t ky = (j < Ny / 2 || (Ny % 2 == 1 && j == Ny / 2)) ? j : j - Ny;
        for (int64_t k = 0; k < Nzh; ++k) sv[(size_t)(j * Nzh + k)] = ky * ky + k * k;
    }
    std::vector<int64_t> us(sv);
    std::sort(us.begin(), us.end());
    us.erase(std::unique(us.begin(), us.end()), us.end());
    L->nu = (int64_t)us.size();
    L->off_s = round256(L->nl * 4);
    L->off_bs = L->off_s + round256(L->nu * 8);
    L->bytes = L->off_bs + L->nu * 2 * 16 * (int64_t)sizeof(double4);
    if (map) {
        map->resize((size_t)L->nl);
        for (size_t i = 0; i < sv.size(); ++i)
            (*map)[i] = (int)(std::lower_bound(us.begin(), us.end(), sv[i]) - us.begin());
    }
    if (svals) {
        svals->resize(us.size());
        for (size_t i = 0; i < us.size(); ++i) (*svals)[i] = (double)us[i];
    }
}

// The plan's line map and distinct s = ky^2 + kz^2 depend on (Ny, Nz) only:
// per device and (Ny, Nz) they are built on the host and uploaded once (the
// first plan of a shape synchronises), then every plan copies them on its
// stream, so that planning a new alpha is asynchronous.
struct PlanHead {
    void* d;        // [linemap | s] as in the plan, off_bs bytes
    PlanLayout L;   // layout at Nx = 8 (E, K and bytes are recomputed per Nx)
};
static std::mutex g_head_mu;
static std::map<std::tuple<int, int64_t, int64_t>, PlanHead> g_heads;

static int plan_head(int64_t Ny, int64_t Nz, PlanHead* out)
{
    int dev = 0;
    SFA_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_head_mu);
    const auto key = std::make_tuple(dev, Ny, Nz);
    auto it = g_heads.find(key);
    if (it != g_heads.end()) {
        *out = it->second;
        return SFA_OK;
    }
    SFA_CHECK(!alloc_forbidden(), SFA_ERR_ARG, "plan head of %lld x %lld was not built",
              (long long)Ny, (long long)Nz);
    PlanHead h;
    std::vector<int> map;
    std::vector<double> sv;
    plan_shape(8, Ny, Nz, &h.L, &map, &sv);
    std::vector<char> hb((size_t)h.L.off_bs, 0);
    memcpy(hb.data(), map.data(), map.size() * sizeof(int));
    memcpy(hb.data() + h.L.off_s, sv.data(), sv.size() * sizeof(double));
    SFA_HIP(hipMalloc(&h.d, hb.size()));
    const hipError_t e = hipMemcpy(h.d, hb.data(), hb.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(h.d);
        return fail(SFA_ERR_HIP, "plan head upload failed: %s", hipGetErrorString(e));
    }
    g_heads[key] = h;
    *out = h;
    return SFA_OK;
}

struct PlanInfo {
    int64_t Nx, Ny, Nz, bytes;
    double alpha;
    int neumann;
    int64_t off_s, off_bs;
    const double* lift;  // added to every solution of this plan (NULL: none)
};
static std::mutex g_plan_mu;
static std::map<const void*, PlanInfo> g_plans;

// Device view of a registered plan; false if the plan does not match (sizes,
// alpha, bc) or was not built by sfa_helmholtz3d_plan.
static bool plan_view(const void* plan, int64_t Nx, int64_t Ny, int64_t Nz, double alpha, int neumann,
                      XPlan* v)
{
    int64_t off_s, off_bs;
    {
        std::lock_guard<std::mutex> lk(g_plan_mu);
        auto it = g_plans.find(plan);
        if (it == g_plans.end()) return false;
        const PlanInfo& p = it->second;
        if (!(p.Nx == Nx && p.Ny == Ny && p.Nz == Nz && p.alpha == alpha && p.neumann == neumann))
            return false;
        off_s = p.off_s;
        off_bs = p.off_bs;
    }
    if (v) {
        const char* pb = (const char*)plan;
        v->linemap = (const int*)pb;
        v->svals = (const double*)(pb + off_s);
        v->bs = (const double4*)(pb + off_bs);
    }
    return true;
}

static bool plan_ok(const void* plan, int64_t Nx, int64_t Ny, int64_t Nz, double alpha, int neumann)
{
    return plan_view(plan, Nx, Ny, Nz, alpha, neumann, nullptr);
}

// The lift registered with a plan (sfa_helmholtz3d_set_lift), NULL if none.
static const double* plan_lift(const void* plan)
{
    if (plan == nullptr) return nullptr;
    std::lock_guard<std::mutex> lk(g_plan_mu);
    auto it = g_plans.find(plan);
    return it == g_plans.end() ? nullptr : it->second.lift;
}

// W: nb arrays (Nx, Ny, Nzh) of complex lines at kz pitch Kp >= Nzh, back to back.
int xpass_launch(double2* W, int Nx, int Ny, int Nzh, int Kp, double alpha, int neumann,
                 hipStream_t s, const void* plan, int64_t nb)
{
    if (plan != nullptr) {
        const double* K2 = cheb_k2(Nx, neumann);
        SFA_CHECK(K2, SFA_ERR_HIP, "table allocation failed");
        XPlan pv;
        SFA_CHECK(plan_view(plan, Nx, Ny, 2 * (Nzh - 1), alpha, neumann, &pv), SFA_ERR_ARG,
                  "plan does not match the solve");
        int rc = 0;
        if (fast_xpass(W, Nx, Ny, Nzh, Kp, alpha, neumann, K2, pv, nb, s, &rc)) return rc;
    }
    if (nb > 1) {  // the generic kernel takes one array per launch
        for (int64_t b = 0; b < nb; ++b) {
            const int rc = xpass_launch(W + b * (int64_t)Nx * Ny * Kp, Nx, Ny, Nzh, Kp, alpha, neumann,
                                        s, nullptr, 1);
            if (rc) return rc;
        }
        return SFA_OK;
    }
    XPassArgs a;
    a.W = W;
    a.sa = (int64_t)Ny * Kp;
    a.s1 = Kp;
    a.d1 = Ny;
    a.d2 = Nzh;
    a.N = Nx;
    a.ldn = Nx + 2;
    // 256 threads = 8 lines x 2 parity chains x 16 lanes
    int T = 8;
    if (T > Nzh) T = Nzh;
    a.T = T;
    const int K = (Nx - 2 + 1) / 2;  // rows of the longer parity chain
    int E = 1;
    while (16 * E < K) E *= 2;
    SFA_CHECK(E <= 16, SFA_ERR_UNSUPPORTED, "Nx=%d too large for the fused x-pass (max 514)", Nx);
    a.E = E;
    a.neumann = neumann;
    a.alpha = alpha;
    a.Ny = Ny;
    a.K2 = cheb_k2(Nx, neumann);
    a.tabs = helm_tables(Nx, neumann);
    a.tw = fft_table(Nx);
    a.tq = quarter_table(Nx);
    SFA_CHECK(a.K2 && a.tabs && a.tw && a.tq, SFA_ERR_HIP, "table allocation failed");
    a.plan = make_fft_plan(Nx);
    const size_t lds = (size_t)2 * T * a.ldn * sizeof(double2) + (size_t)2 * T * sizeof(double);
    SFA_CHECK(lds <= 160 * 1024, SFA_ERR_UNSUPPORTED, "Nx=%d too large for the fused x-pass", Nx);
    if (lds > 64 * 1024)
        hipFuncSetAttribute((const void*)xpass_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lds);
    const int64_t nblk = (int64_t)Ny * ((Nzh + T - 1) / T);
    hipLaunchKernelGGL(xpass_kernel, dim3((unsigned)nblk), dim3(256), lds, s, a);
    SFA_LAUNCH_CHECK();
    return SFA_OK;
}

// y C2C of both sub-arrays of the mixed workspace: one launch whose first
// blocks take the complex128 planes (running the two on two streams measured
// slower); two launches for the lengths the dual kernel does not cover.
// out.flag != NULL: also deliver the statistics in stats (a block of the
// one launch, else a kernel of its own first).
static int y_pass_mixed(const Geom& g64, const Geom& g32, double2* W64, float2* W32, int sign,
                        double scale, hipStream_t s, const double* stats = nullptr,
                        StatOut out = StatOut{nullptr, nullptr, 0u})
{
    int rc = 0;
    if (fast_c2c_mixed(g64, W64, g32, W32, sign, scale, s, &rc, stats, out)) return rc;
    if (out.flag) {
        hipLaunchKernelGGL(stat_deliver_kernel, dim3(1), dim3(64), 0, s, stats, out);
        SFA_LAUNCH_CHECK();
    }
    if (!fast_c2c(g64, g64, W64, W64, sign, scale, s, &rc) || rc ||
        !fast_c2c_f32(g32, g32, W32, W32, sign, scale, s, &rc) || rc)
        return rc ? rc : fail(SFA_ERR_UNSUPPORTED, "mixed precision: Ny=%lld unsupported", (long long)g64.n);
    return SFA_OK;
}

// Mixed-precision pipeline: kz planes [0, Kh) stay complex128 with FP64
// transforms, planes [Kh, Nz/2] are stored and transformed as complex64; the
// first (R2C of the FP64 f) and last (C2R to the FP64 u) transforms and the
// Helmholtz solve are FP64.
// nb arrays (Nx, Ny, Nz) back to back in f and u; the workspace holds the nb
// arrays' complex128 planes, then their complex64 planes.
// stats (may be NULL): band-energy accumulators (kStatBytes), cleared by the z
// R2C split and then added to by the x-pass; out.flag (may be NULL): where the
// inverse y pass delivers the summed accumulators.
static int helmholtz3d_mixed(const double* f, double* u, int64_t Nx, int64_t Ny, int64_t Nz,
                             double alpha, int neumann, int Kh, const void* plan, void* work,
                             hipStream_t s, hipEvent_t* ev, int64_t nb = 1, double* stats = nullptr,
                             StatOut out = StatOut{nullptr, nullptr, 0u})
{
    auto mark = [&](int i) {
        if (ev) hipEventRecord(ev[i], s);
    };
    const int64_t Nzh = Nz / 2 + 1;
    const int Kl = (int)Nzh - Kh;
    XPlan pv;
    SFA_CHECK(plan_view(plan, Nx, Ny, Nz, alpha, neumann, &pv), SFA_ERR_ARG,
              "mixed precision needs a matching plan");
    const double* K2 = cheb_k2((int)Nx, neumann);
    SFA_CHECK(K2, SFA_ERR_HIP, "table allocation failed");
    double2* W64 = (double2*)work;
    float2* W32 = (float2*)(W64 + nb * Nx * Ny * Kh);
    const int64_t rows = nb * Nx * Ny;
    int rc = 0;
    auto geom = [&](int Kp) {
        Geom g;
        g.n = Ny;
        g.sa = Kp;
        g.d1 = nb * Nx;
        g.s1 = Ny * Kp;
        g.d2 = Kp;
        g.s2 = 1;
        return g;
    };
    const Geom g64 = geom(Kh), g32 = geom(Kl);
    mark(0);
    // the split also clears the spectrum statistics that the x-pass adds into
    if (!split_rc(true, (int)(Nz / 2), f, W64, W32, rows, Kh, 1.0 / (double)Nz, s, &rc, stats,
                  stats ? kStatBytes / 8 : 0))
        return fail(SFA_ERR_UNSUPPORTED, "mixed precision: Nz=%lld unsupported", (long long)Nz);
    if (rc) return rc;
    mark(1);
    if ((rc = y_pass_mixed(g64, g32, W64, W32, -1, 1.0 / (double)Ny, s))) return rc;
    mark(2);
    if (!fast_xpass_mixed(W64, W32, (int)Nx, (int)Ny, (int)Nzh, Kh, alpha, neumann, K2, pv, nb, s,
                          &rc, stats) ||
        rc)
        return rc ? rc : fail(SFA_ERR_UNSUPPORTED, "mixed precision: Nx=%lld unsupported", (long long)Nx);
    mark(3);
    if ((rc = y_pass_mixed(g64, g32, W64, W32, +1, 1.0, s, stats, out))) return rc;
    mark(4);
    // the plan's lift (nonhomogeneous boundary values) is added by the C2R's stores
    const double* lift = plan_lift(plan);
    if (!split_rc(false, (int)(Nz / 2), u, W64, W32, rows, Kh, 1.0, s, &rc, nullptr, 0, lift, Nx * Ny))
        return fail(SFA_ERR_UNSUPPORTED, "mixed precision: Nz unsupported");
    mark(5);
    return rc;
}

// kz pitch of the FP64 workspace: Nzh rounded up to a multiple of 8 complex
// (128 bytes) where the fast x-pass runs and the padding is under 1/8 of the
// planes; otherwise Nzh.
static int64_t fp64_pitch(int64_t Nx, int64_t Nzh, const void* plan)
{
    const bool fast = plan != nullptr && (Nx == 64 || Nx == 128 || Nx == 256 || Nx == 512);
    return (fast && Nzh >= 64) ? (Nzh + 7) / 8 * 8 : Nzh;
}

// FP64 pipeline: z R2C, y C2C, fused x-pass, y C2C inverse, z C2R.  ev: NULL
// or 6 events recorded around the five stages.  nb arrays (Nx, Ny, Nz) back
// to back in f, u and (at the padded pitch) the workspace: the z and y passes
// see one (nb Nx, Ny, .) array, the x-pass offsets each array's lines.
static int helmholtz3d_fp64(const double* f, double* u, int64_t Nx, int64_t Ny, int64_t Nz,
                            double alpha, int neumann, const void* plan, void* work, hipStream_t s,
                            hipEvent_t* ev, int64_t nb = 1)
{
    auto mark = [&](int i) {
        if (ev) hipEventRecord(ev[i], s);
    };
    const int64_t Nzh = Nz / 2 + 1;
    const int64_t Kp = fp64_pitch(Nx, Nzh, plan);
    double2* W = (double2*)work;
    const int64_t rshape[3] = {nb * Nx, Ny, Nz};
    const int64_t rstr[3] = {Ny * Nz, Nz, 1};
    const int64_t cstr[3] = {Ny * Kp, Kp, 1};
    // the y passes run over the padding lines too, so that each workgroup's
    // 16-line row segments start on 128-byte boundaries
    const int64_t yshape[3] = {nb * Nx, Ny, Kp};
    void* st = (void*)s;
    int rc;
    mark(0);
    if ((rc = sfa_fft_r2c(f, W, 3, rshape, rstr, cstr, 2, 1.0 / (double)Nz, st))) return rc;
    mark(1);
    if ((rc = sfa_fft_c2c(W, W, 3, yshape, cstr, cstr, 1, -1, 1.0 / (double)Ny, st))) return rc;
    mark(2);
    if ((rc = xpass_launch(W, (int)Nx, (int)Ny, (int)Nzh, (int)Kp, alpha, neumann, s, plan, nb)))
        return rc;
    mark(3);
    if ((rc = sfa_fft_c2c(W, W, 3, yshape, cstr, cstr, 1, +1, 1.0, st))) return rc;
    mark(4);
    // the plan's lift (nonhomogeneous boundary values) is added by the C2R
    if ((rc = fft_c2r_add(W, u, 3, rshape, cstr, rstr, 2, 1.0, st, plan_lift(plan), Nx * Ny))) return rc;
    mark(5);
    return SFA_OK;
}

static int mixed_planes(int64_t Nz, int fp64_planes)
{
    const int64_t Nzh = Nz / 2 + 1;
    // 17: the complex64 planes' pitch Nz/2 + 1 - 17 is then a multiple of 16
    // (128-byte rows) for every power-of-two Nz >= 32, which the y passes need
    // for aligned row segments (Kh = 16 measured 1.35x slower y passes at 256^3)
    int Kh = fp64_planes > 0 ? fp64_planes : 17;
    if (Kh < 2) Kh = 2;
    if (Kh > Nzh - 2) Kh = (int)Nzh - 2;
    return Kh;
}

}  // namespace sfa

using namespace sfa;

// Complex workspace of one array (FP64 layout at its padded pitch; the mixed
// layout is smaller); the guarded solve keeps its statistics behind the nb
// arrays' workspaces.
static int64_t core_workspace(int64_t Nx, int64_t Ny, int64_t Nz)
{
    const int64_t Nzh = Nz / 2 + 1;
    const int64_t Kp = Nzh >= 64 ? (Nzh + 7) / 8 * 8 : Nzh;  // covers fp64_pitch()
    return Nx * Ny * Kp * (int64_t)sizeof(double2);
}

extern "C" int64_t sfa_helmholtz3d_workspace(int64_t Nx, int64_t Ny, int64_t Nz, int precision)
{
    (void)precision;
    return core_workspace(Nx, Ny, Nz) + kStatBytes;
}

extern "C" int64_t sfa_helmholtz3d_plan_bytes(int64_t Nx, int64_t Ny, int64_t Nz)
{
    if (Nx < 8 || Ny < 1 || Nz < 2) return -1;
    PlanLayout L;
    plan_shape(Nx, Ny, Nz, &L, nullptr, nullptr);
    return L.bytes;
}

extern "C" int sfa_helmholtz3d_plan(int64_t Nx, int64_t Ny, int64_t Nz, double alpha, int neumann,
                                    void* plan, int64_t plan_bytes, void* stream)
{
    SFA_CHECK(Nx >= 8 && Ny >= 1 && Nz >= 2 && Nz % 2 == 0, SFA_ERR_ARG,
              "need Nx >= 8, Ny >= 1, even Nz >= 2");
    SFA_CHECK(plan != nullptr, SFA_ERR_ARG, "plan buffer is NULL");
    PlanHead head;
    int rc = plan_head(Ny, Nz, &head);
    if (rc) return rc;
    PlanLayout L = head.L;
    L.K = (Nx - 2 + 1) / 2;
    L.E = (L.K + 15) / 16;
    SFA_CHECK(plan_bytes >= L.bytes, SFA_ERR_ARG, "plan buffer too small (%lld < %lld)",
              (long long)plan_bytes, (long long)L.bytes);
    hipStream_t s = (hipStream_t)stream;
    char* pb = (char*)plan;
    SFA_HIP(hipMemcpyAsync(pb, head.d, (size_t)L.off_bs, hipMemcpyDeviceToDevice, s));
    const int M = (int)Nx - 2;
    const int64_t nth = 2 * L.nu;
    hipLaunchKernelGGL(helm_bstate_kernel, dim3((unsigned)((nth + 127) / 128)), dim3(128), 0, s, M,
                       (int)L.E, neumann, alpha, (const double*)(pb + L.off_s), L.nu,
                       (double4*)(pb + L.off_bs));
    SFA_LAUNCH_CHECK();
    // the device tables of the pipeline's kernels are built here, at set-up,
    // so that the no-allocation solve (which only looks them up) finds them
    SFA_CHECK(fft_table((int)Nx) && quarter_table((int)Nx) && fft_table((int)Ny) &&
                  fft_table((int)Nz) && cheb_k2((int)Nx, neumann),
              SFA_ERR_HIP, "table allocation failed");
    {
        std::lock_guard<std::mutex> lk(g_plan_mu);
        g_plans[plan] = PlanInfo{Nx, Ny, Nz, plan_bytes, alpha, neumann, L.off_s, L.off_bs, nullptr};
    }
    return SFA_OK;
}

extern "C" int sfa_helmholtz3d_set_lift(const void* plan, const double* lift)
{
    SFA_CHECK(lift == nullptr || (reinterpret_cast<uintptr_t>(lift) & 15) == 0, SFA_ERR_ARG,
              "lift must be 16-byte aligned");
    std::lock_guard<std::mutex> lk(g_plan_mu);
    auto it = g_plans.find(plan);
    SFA_CHECK(it != g_plans.end(), SFA_ERR_ARG, "plan was not built by sfa_helmholtz3d_plan");
    it->second.lift = lift;
    return SFA_OK;
}

extern "C" int sfa_helmholtz3d_plan_release(const void* plan)
{
    std::lock_guard<std::mutex> lk(g_plan_mu);
    g_plans.erase(plan);
    return SFA_OK;
}

extern "C" int sfa_helmholtz3d_many(const double* f, double* u, int64_t nbatch, int64_t Nx,
                                    int64_t Ny, int64_t Nz, double alpha, int neumann,
                                    int precision, int fp64_planes, const void* plan, void* work,
                                    int64_t work_bytes, void* stream)
{
    SFA_CHECK(plan == nullptr || plan_ok(plan, Nx, Ny, Nz, alpha, neumann), SFA_ERR_ARG,
              "plan was not built by sfa_helmholtz3d_plan for these sizes/alpha/bc");
    SFA_CHECK(Nx >= 8 && Ny >= 1 && Nz >= 2 && Nz % 2 == 0, SFA_ERR_ARG,
              "need Nx >= 8, Ny >= 1, even Nz >= 2");
    SFA_CHECK(nbatch >= 1, SFA_ERR_ARG, "nbatch must be >= 1");
    SFA_CHECK(f != nullptr && u != nullptr, SFA_ERR_ARG, "NULL f or u");
    SFA_CHECK(precision == 0 || precision == 1, SFA_ERR_ARG, "precision must be 0 or 1");
    SFA_CHECK(work_bytes / nbatch >= core_workspace(Nx, Ny, Nz), SFA_ERR_ARG, "workspace too small");
    hipStream_t s = (hipStream_t)stream;
    if (precision == 1) {
        SFA_CHECK(Nz / 2 + 1 >= 4, SFA_ERR_UNSUPPORTED, "mixed precision needs Nz >= 6");
        return helmholtz3d_mixed(f, u, Nx, Ny, Nz, alpha, neumann, mixed_planes(Nz, fp64_planes),
                                 plan, work, s, nullptr, nbatch);
    }
    return helmholtz3d_fp64(f, u, Nx, Ny, Nz, alpha, neumann, plan, work, s, nullptr, nbatch);
}

extern "C" int sfa_helmholtz3d(const double* f, double* u, int64_t Nx, int64_t Ny, int64_t Nz,
                               double alpha, int neumann, int precision, int fp64_planes,
                               const void* plan, void* work, int64_t work_bytes, void* stream)
{
    return sfa_helmholtz3d_many(f, u, 1, Nx, Ny, Nz, alpha, neumann, precision, fp64_planes, plan,
                                work, work_bytes, stream);
}

// ------------------------------------------------------------- no-allocation solve
// The workspace of sfa_helmholtz3d_noalloc: one file-scope device array per
// GPU, held until the process ends.  Host code never passes its name to a
// kernel (in host code the name is a host-side stand-in, not a device
// address): the device address comes from hipGetSymbolAddress, once per
// device, and only that address is used.
static constexpr int64_t kNoAllocScratchBytes = 32ll << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

static std::mutex g_scratch_mu;
static std::map<int, void*> g_scratch_ptr;  // device -> address of its `scratch`

static void* scratch_address(int* rc)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        *rc = fail(SFA_ERR_HIP, "hipGetDevice failed: %s", hipGetErrorString(e));
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_ptr.find(dev);
    if (it != g_scratch_ptr.end()) return it->second;
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
        *rc = fail(SFA_ERR_HIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                   hipGetErrorString(e));
        return nullptr;
    }
    g_scratch_ptr[dev] = p;
    return p;
}

extern "C" int64_t sfa_helmholtz3d_noalloc_max_bytes(void) { return kNoAllocScratchBytes; }

extern "C" int sfa_helmholtz3d_noalloc(const double* f, double* u, int64_t Nx, int64_t Ny,
                                       int64_t Nz, double alpha, int neumann, int precision,
                                       int fp64_planes, const void* plan, void* stream)
{
    SFA_CHECK(Nx >= 8 && Ny >= 1 && Nz >= 2 && Nz % 2 == 0, SFA_ERR_ARG,
              "need Nx >= 8, Ny >= 1, even Nz >= 2");
    SFA_CHECK(f != nullptr && u != nullptr, SFA_ERR_ARG, "NULL f or u");
    SFA_CHECK(precision == 0 || precision == 1, SFA_ERR_ARG, "precision must be 0 or 1");
    SFA_CHECK(plan != nullptr && plan_ok(plan, Nx, Ny, Nz, alpha, neumann), SFA_ERR_ARG,
              "the no-allocation solve needs a plan built by sfa_helmholtz3d_plan for these "
              "sizes/alpha/bc");
    auto pow2 = [](int64_t n) { return n > 0 && (n & (n - 1)) == 0; };
    SFA_CHECK((Nx == 64 || Nx == 128 || Nx == 256 || Nx == 512) && pow2(Ny) && Ny >= 16 &&
                  Ny <= 512 && pow2(Nz) && Nz >= 32 && Nz <= 512,
              SFA_ERR_UNSUPPORTED,
              "the no-allocation solve takes Nx in {64, 128, 256, 512} and power-of-two Ny in "
              "[16, 512], Nz in [32, 512]");
    const int64_t need = core_workspace(Nx, Ny, Nz);
    SFA_CHECK(need <= kNoAllocScratchBytes, SFA_ERR_ARG,
              "%lld x %lld x %lld needs %lld bytes of workspace; the no-allocation solve holds "
              "%lld bytes (use sfa_helmholtz3d with a caller workspace)",
              (long long)Nx, (long long)Ny, (long long)Nz, (long long)need,
              (long long)kNoAllocScratchBytes);
    int rc = SFA_OK;
    void* w = scratch_address(&rc);
    if (w == nullptr) return rc;
    NoAllocScope guard;
    // every table the kernels read was built by sfa_helmholtz3d_plan
    SFA_CHECK(fft_table((int)Nx) && quarter_table((int)Nx) && fft_table((int)Ny) &&
                  fft_table((int)Nz) && cheb_k2((int)Nx, neumann),
              SFA_ERR_ARG, "device tables missing: call sfa_helmholtz3d_plan on this device first");
    hipStream_t s = (hipStream_t)stream;
    if (precision == 1)
        return helmholtz3d_mixed(f, u, Nx, Ny, Nz, alpha, neumann, mixed_planes(Nz, fp64_planes),
                                 plan, w, s, nullptr, 1);
    return helmholtz3d_fp64(f, u, Nx, Ny, Nz, alpha, neumann, plan, w, s, nullptr, 1);
}

namespace sfa {

// Error model of the mixed form.  Complex64 storage and FP32 transforms of
// the planes kz >= Kh add relative rounding errors of a few eps32 per stage to
// those planes' values: forward (the stored z spectrum of f, the y FFT, the
// DCT-II) before the solve, whose gain on those planes is at most
// 1 / (alpha + Kh^2), and backward (DCT-III, inverse y FFT, the stored
// spectrum) on the solution's own values.  Rounding errors add like noise, so
// the rms error of u is about eps32 sqrt(w^2 E_f,hi + E_u,hi) with E the grid
// energies of f and u in those planes (Parseval: E_f = Ny Nz sum w_kz |F_x|^2
// over the x-pass input lines F (kz, ky, x), the z spectrum scaled by 1/Nz and
// the y transform by 1/Ny; E_u = Nx Ny Nz sum w_kz (|t_0|^2 + 1/2 sum |t_k|^2)
// over u's Chebyshev coefficients t, the DCT-II being orthogonal on the
// Gauss-Chebyshev points), and max|u| >= rms(u) = sqrt(E_u / (Nx Ny Nz)).
// kGuardK (64) covers the number of rounding stages (about 30, adding like a
// random walk) and the ratio of the largest to the rms error.  Bands
// straddling Kh count as complex64 (conservative).
constexpr double kEps32 = 5.9604644775390625e-08;
constexpr double kGuardK = 64.0;

static double guard_predict(const double* h, int Kh, int64_t Nx, int64_t Ny, int64_t Nz, double alpha)
{
    const double* sf = h;  // the reduced sums: f's bands, then u's
    const double* st = h + kStatBands;
    const int b0 = stat_band(Kh);
    double ef_hi = 0.0, eu_hi = 0.0, eu = 0.0;
    for (int b = 0; b < kStatBands; ++b) {
        const double ef = sf[b], et = st[b];
        eu += et;
        if (b >= b0) {
            ef_hi += ef;
            eu_hi += et;
        }
    }
    const double vol = (double)Nx * (double)Ny * (double)Nz;
    ef_hi *= (double)Ny * (double)Nz;
    eu_hi *= vol;
    eu *= vol;
    const double den = alpha + (double)Kh * (double)Kh;
    const double w = den > 1.0 ? 1.0 / den : 1.0;
    if (!(eu > 0.0)) return (ef_hi + eu_hi > 0.0) ? INFINITY : 0.0;
    const double r = (w * w * ef_hi + eu_hi) / eu;
    return r >= 0.0 ? kGuardK * kEps32 * sqrt(r) : INFINITY;
}

// Per host thread and device: a pinned, device-mapped, fine-grained host
// buffer into which the inverse y pass writes the summed statistics and then
// the call's sequence number.  It lives until the process ends.
struct GuardHost {
    double sums[kStatAcc];
    volatile unsigned flag;
};

struct GuardCtx {
    GuardHost* host;
    GuardHost* dev;  // device address of host
    unsigned seq;
};

static GuardCtx* guard_ctx()
{
    static thread_local std::map<int, GuardCtx> ctxs;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    auto it = ctxs.find(dev);
    if (it != ctxs.end()) return &it->second;
    GuardCtx c{};
    if (hipHostMalloc((void**)&c.host, sizeof(GuardHost), hipHostMallocMapped | hipHostMallocCoherent) !=
        hipSuccess)
        return nullptr;
    if (hipHostGetDevicePointer((void**)&c.dev, c.host, 0) != hipSuccess) {
        hipHostFree(c.host);
        return nullptr;
    }
    c.host->flag = 0u;
    return &(ctxs[dev] = c);
}

// Spin until call seq's sums are published.  Once per millisecond of waiting
// the stream is queried (a query can append a marker to the stream's queue,
// which measured as a ~6 us stall before the next call's first kernel when
// issued every few microseconds): an error ends the wait with that error, and
// a stream that has drained without the flag set is reported as a failure.
static int guard_wait(const GuardCtx* g, unsigned seq, hipStream_t s)
{
    using clk = std::chrono::steady_clock;
    auto next = clk::now() + std::chrono::milliseconds(1);
    for (unsigned it = 1;; ++it) {
        if (g->host->flag == seq) break;
        __builtin_ia32_pause();  // spin-wait hint: yields the core's pipeline to a sibling thread
        if ((it & 1023u) == 0 && clk::now() >= next) {
            const hipError_t q = hipStreamQuery(s);
            if (q == hipSuccess) {
                if (g->host->flag == seq) break;
                return fail(SFA_ERR_HIP, "guarded solve: statistics were not delivered");
            }
            if (q != hipErrorNotReady) SFA_HIP(q);
            next = clk::now() + std::chrono::milliseconds(1);
        }
    }
    __atomic_thread_fence(__ATOMIC_ACQUIRE);  // the sums were written before the flag
    return SFA_OK;
}

}  // namespace sfa

extern "C" int sfa_helmholtz3d_guarded(const double* f, double* u, int64_t nbatch, int64_t Nx,
                                       int64_t Ny, int64_t Nz, double alpha, int neumann,
                                       int kh_hint, double tol, const void* plan, void* work,
                                       int64_t work_bytes, int* kh_used, int* kh_next,
                                       double* err_pred, void* stream)
{
    SFA_CHECK(Nx >= 8 && Ny >= 1 && Nz >= 2 && Nz % 2 == 0, SFA_ERR_ARG,
              "need Nx >= 8, Ny >= 1, even Nz >= 2");
    SFA_CHECK(plan != nullptr && plan_ok(plan, Nx, Ny, Nz, alpha, neumann), SFA_ERR_ARG,
              "plan was not built by sfa_helmholtz3d_plan for these sizes/alpha/bc");
    SFA_CHECK(nbatch >= 1, SFA_ERR_ARG, "nbatch must be >= 1");
    SFA_CHECK(f != nullptr && u != nullptr && work != nullptr, SFA_ERR_ARG, "NULL f, u or work");
    SFA_CHECK(tol > 0.0, SFA_ERR_ARG, "tol must be > 0");
    const int64_t Nzh = Nz / 2 + 1;
    SFA_CHECK(Nzh >= 4, SFA_ERR_UNSUPPORTED, "mixed precision needs Nz >= 6");
    SFA_CHECK(Nz <= 2 * 16 * kStatBands, SFA_ERR_UNSUPPORTED, "guarded mixed precision: Nz <= 512");
    const int64_t core = core_workspace(Nx, Ny, Nz);
    SFA_CHECK(work_bytes >= nbatch * core + kStatBytes, SFA_ERR_ARG,
              "workspace too small (need nbatch * %lld + %d bytes)", (long long)core, kStatBytes);
    hipStream_t s = (hipStream_t)stream;
    GuardCtx* g = guard_ctx();
    SFA_CHECK(g != nullptr, SFA_ERR_HIP, "guard pinned buffer creation failed");
    double* stats = (double*)((char*)work + nbatch * core);  // cleared by the z R2C split
    const int Kh = mixed_planes(Nz, kh_hint);
    const unsigned seq = ++g->seq == 0u ? ++g->seq : g->seq;  // 0 is the buffer's initial flag
    // The statistics are complete when the x-pass ends; the first workgroup
    // of the inverse y pass sums them into the host buffer, and the host's
    // wait for that overlaps the inverse y and z passes.  The wait spins: the next call must be
    // enqueued before those ~70 us (256^3) run out, and a blocking wait's
    // wake-up latency had let the device idle between solves.
    int rc = helmholtz3d_mixed(f, u, Nx, Ny, Nz, alpha, neumann, Kh, plan, work, s, nullptr, nbatch,
                               stats, StatOut{g->dev->sums, const_cast<unsigned*>(&g->dev->flag), seq});
    if (rc) return rc;
    if ((rc = guard_wait(g, seq, s))) return rc;
    double e = guard_predict(g->host->sums, Kh, Nx, Ny, Nz, alpha);
    // admissible first complex64 planes 16m + 1 (the band edges)
    int next = 0;
    double enext = 0.0;
    for (int m = 1; 16 * m + 1 <= Nzh - 2; ++m) {
        const double em = guard_predict(g->host->sums, 16 * m + 1, Nx, Ny, Nz, alpha);
        if (em <= tol) {
            next = 16 * m + 1;
            enext = em;
            break;
        }
    }
    int used = Kh;
    if (!(e <= tol)) {
        // redo from f: with more complex128 planes, or in FP64
        if (next > Kh) {
            used = next;
            e = enext;
            rc = helmholtz3d_mixed(f, u, Nx, Ny, Nz, alpha, neumann, used, plan, work, s, nullptr, nbatch);
        } else {
            used = (int)Nzh;
            e = 0.0;
            rc = helmholtz3d_fp64(f, u, Nx, Ny, Nz, alpha, neumann, plan, work, s, nullptr, nbatch);
        }
        if (rc) return rc;
    }
    if (kh_used) *kh_used = used;
    if (kh_next) *kh_next = next;
    if (err_pred) *err_pred = e;
    return SFA_OK;
}

// Same pipeline as sfa_helmholtz3d, run nrep times with hipEvents recorded on
// `stream` between the five kernels; stage_ms[5] receives the summed
// milliseconds of each stage (z R2C, y C2C, x-pass, y C2C inverse, z C2R).
extern "C" int sfa_helmholtz3d_stage_times(const double* f, double* u, int64_t Nx, int64_t Ny,
                                           int64_t Nz, double alpha, int neumann, int precision,
                                           int fp64_planes, const void* plan, void* work,
                                           int64_t work_bytes, int nrep, float* stage_ms,
                                           void* stream)
{
    SFA_CHECK(plan == nullptr || plan_ok(plan, Nx, Ny, Nz, alpha, neumann), SFA_ERR_ARG,
              "plan was not built by sfa_helmholtz3d_plan for these sizes/alpha/bc");
    SFA_CHECK(Nx >= 8 && Ny >= 1 && Nz >= 2 && Nz % 2 == 0, SFA_ERR_ARG,
              "need Nx >= 8, Ny >= 1, even Nz >= 2");
    SFA_CHECK(precision >= 0 && precision <= 2, SFA_ERR_ARG, "precision must be 0, 1 or 2");
    SFA_CHECK(precision == 0 || Nz / 2 + 1 >= 4, SFA_ERR_UNSUPPORTED, "mixed precision needs Nz >= 6");
    SFA_CHECK(work_bytes >= sfa_helmholtz3d_workspace(Nx, Ny, Nz, precision), SFA_ERR_ARG,
              "workspace too small");
    double* stats = precision == 2 ? (double*)((char*)work + core_workspace(Nx, Ny, Nz)) : nullptr;
    SFA_CHECK(nrep >= 1 && stage_ms != nullptr, SFA_ERR_ARG, "bad nrep/stage_ms");
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t ev[6];
    for (int i = 0; i < 6; ++i) SFA_HIP(hipEventCreate(&ev[i]));
    for (int i = 0; i < 5; ++i) stage_ms[i] = 0.f;
    int rc = SFA_OK;
    for (int r = 0; r < nrep && rc == SFA_OK; ++r) {
        rc = precision >= 1 ? helmholtz3d_mixed(f, u, Nx, Ny, Nz, alpha, neumann,
                                                mixed_planes(Nz, fp64_planes), plan, work, s, ev, 1,
                                                stats)
                            : helmholtz3d_fp64(f, u, Nx, Ny, Nz, alpha, neumann, plan, work, s, ev);
        if (rc) break;
        hipEventSynchronize(ev[5]);
        for (int i = 0; i < 5; ++i) {
            float ms = 0.f;
            hipEventElapsedTime(&ms, ev[i], ev[i + 1]);
            stage_ms[i] += ms;
        }
    }
    hipStreamSynchronize(s);
    for (int i = 0; i < 6; ++i) hipEventDestroy(ev[i]);
    return rc;
}
