// Wavefront-parallel Chebyshev Helmholtz LU + solve of one parity chain.
//
// The reference factorises alfa*A + beta*B row by row (LU_Helmholtz_1D,
// la.pyx:717-763) and solves with a forward sweep and a back substitution
// carrying running parity sums (Solve_Helmholtz_1D_ptr, la.pyx:871-912).
// All band offsets are even, so the even and odd unknowns form two
// independent chains of K <= M/2 + 1 rows.  Here one chain is handled by 16
// lanes (E consecutive rows each) with three scans across the 16 lanes:
//
//  1. LU: in homogeneous coordinates the row recurrence
//       d0 = h - g d1'/d0', d1 = s - g d2'/d0', d2 = c - g d2'/d0'
//     is linear: zeta_m = N_m zeta_{m-1}, N_m = [[h,-g,0],[s,0,-g],[c,0,-g]],
//     zeta_m proportional to (d0, d1, d2)_m.  Segment products of N are
//     scanned (normalised by powers of two); each lane then restarts the
//     exact row recurrence from the ratios of its exclusive prefix.
//  2. forward elimination y_m = f_m - (g_m/d0_{m-1}) y_{m-1}: affine scan.
//  3. back substitution in the state (v_m, T_m = sum_{j>=m} v_j):
//       v_m = y_m/d0 - ((d1-d2)/d0) v_{m+1} - (d2/d0) T_{m+1},  T_m = v_m + T_{m+1}
//     a 2x2 affine suffix scan.
#pragma once
#include "common.h"

namespace sfa {

// Row coefficients of one chain row (global index i), Neumann column scaling
// and k = 0 gauge of la.pyx:735-749 included.
struct HelmRow {
    double h, s, c, g;
};

struct HelmTab {
    const double *A0, *A2, *A4, *Bm2, *B0, *B2;  // unscaled diagonals, M entries each
};

__device__ inline HelmRow helm_row(const HelmTab& t, int i, int M, double al, double be, bool nm,
                                   bool zb)
{
    HelmRow r;
    if (i >= M) {  // padding row: identity, no coupling
        r.h = 1.0;
        r.s = r.c = r.g = 0.0;
        return r;
    }
    double a0 = t.A0[i], b0 = t.B0[i];
    double a2 = (i < M - 2) ? t.A2[i] : 0.0, b2 = (i < M - 2) ? t.B2[i] : 0.0;
    double a4 = (i < M - 4) ? t.A4[i] : 0.0;
    double bm = (i >= 2) ? t.Bm2[i - 2] : 0.0;
    if (nm) {
        if (i == 0) {
            if (zb) { a0 = 1.0 / al; b0 = 0.0; }
        } else {
            const double ii = 1.0 / ((double)i * i);
            a0 *= ii;
            b0 *= ii;
        }
        const double i2 = 1.0 / ((double)(i + 2) * (i + 2));
        a2 *= i2;
        b2 *= i2;
        a4 *= 1.0 / ((double)(i + 4) * (i + 4));
        if (i - 2 >= 1) bm *= 1.0 / ((double)(i - 2) * (i - 2));
    }
    r.h = al * a0 + be * b0;
    r.s = al * a2 + be * b2;
    r.c = al * a4;
    r.g = be * bm;
    return r;
}

struct M3 {
    double a[9];
};

__device__ inline M3 m3mul(const M3& x, const M3& y)  // x @ y
{
    M3 z;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            z.a[3 * i + j] = x.a[3 * i] * y.a[j] + x.a[3 * i + 1] * y.a[3 + j] + x.a[3 * i + 2] * y.a[6 + j];
    return z;
}

__device__ inline void m3norm(M3& x)
{
    double mx = 0.0;
#pragma unroll
    for (int i = 0; i < 9; ++i) mx = fmax(mx, fabs(x.a[i]));
    if (mx > 0.0) {
        int e;
        frexp(mx, &e);
#pragma unroll
        for (int i = 0; i < 9; ++i) x.a[i] = ldexp(x.a[i], -e);
    }
}

// Lane shifts inside aligned groups of W lanes with __shfl_up/down semantics
// (a lane whose source falls outside its group keeps its own value).  For
// W = 16 a group is one DPP row, so the shift is a DPP row_shr/row_shl move on
// the VALU instead of an LDS-crossbar ds_bpermute (the x-pass is LDS-bound
// too: 6% faster at 256^3).  All 16 lanes of a group are active at every call.
template <int CTRL>
__device__ inline double dpp_row64(double v)
{
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = (int)b, hi = (int)(b >> 32);
    const int nlo = __builtin_amdgcn_update_dpp(lo, lo, CTRL, 0xF, 0xF, false);
    const int nhi = __builtin_amdgcn_update_dpp(hi, hi, CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((long long)nhi << 32) | (unsigned int)nlo);
}
template <int W>
__device__ inline double shup(double v, int d)
{
    if constexpr (W == 16) {
        switch (d) {  // d is a constant after unrolling
            case 1: return dpp_row64<0x111>(v);  // row_shr:1
            case 2: return dpp_row64<0x112>(v);
            case 4: return dpp_row64<0x114>(v);
            case 8: return dpp_row64<0x118>(v);
            default: break;
        }
    }
    return __shfl_up(v, d, W);
}
template <int W>
__device__ inline double shdn(double v, int d)
{
    if constexpr (W == 16) {
        switch (d) {
            case 1: return dpp_row64<0x101>(v);  // row_shl:1
            case 2: return dpp_row64<0x102>(v);
            case 4: return dpp_row64<0x104>(v);
            case 8: return dpp_row64<0x108>(v);
            default: break;
        }
    }
    return __shfl_down(v, d, W);
}

// Solve one chain.  f: right-hand side of this lane's rows i = 2 (lane E + e) + p
// (zero for i >= M); X: the line (LDS), solution rows written there.
// W lanes (aligned group) per chain; lane in [0, W).
template <int E, int W = 16>
__device__ inline void helm_chain_scan_f(const double2 (&f)[E], double2* X, int p, int M, double al,
                                         double be, bool nm, const HelmTab& tab, int lane)
{
    const bool zb = nm && fabs(be) < 1e-8;
    const int m0 = lane * E;
    // ---- 1. LU scan
    M3 seg;
#pragma unroll
    for (int i = 0; i < 9; ++i) seg.a[i] = (i % 4 == 0) ? 1.0 : 0.0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const HelmRow r = helm_row(tab, 2 * (m0 + e) + p, M, al, be, nm, zb);
        M3 Nm;
        Nm.a[0] = r.h; Nm.a[1] = -r.g; Nm.a[2] = 0.0;
        Nm.a[3] = r.s; Nm.a[4] = 0.0;  Nm.a[5] = -r.g;
        Nm.a[6] = r.c; Nm.a[7] = 0.0;  Nm.a[8] = -r.g;
        seg = m3mul(Nm, seg);
    }
    m3norm(seg);
#pragma unroll
    for (int k = 1; k < W; k <<= 1) {
        M3 o;
#pragma unroll
        for (int i = 0; i < 9; ++i) o.a[i] = shup<W>(seg.a[i], k);
        if (lane >= k) {
            seg = m3mul(seg, o);
            m3norm(seg);
        }
    }
    // exclusive prefix applied to zeta_{-1} = (1, 0, 0): first column of lane-1's product
    double z0 = shup<W>(seg.a[0], 1), z1 = shup<W>(seg.a[3], 1), z2 = shup<W>(seg.a[6], 1);
    if (lane == 0) { z0 = 1.0; z1 = 0.0; z2 = 0.0; }
    double r1 = z1 / z0, r2 = z2 / z0;
    double inv[E], e1[E], e2[E], am[E];
    double lastinv = 0.0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const HelmRow r = helm_row(tab, 2 * (m0 + e) + p, M, al, be, nm, zb);
        const double d0 = r.h - r.g * r1;
        const double d1 = r.s - r.g * r2;
        const double d2 = r.c - r.g * r2;
        const double iv = 1.0 / d0;
        am[e] = -r.g;  // times 1/d0_{m-1}, filled below
        inv[e] = iv;
        r1 = d1 * iv;
        r2 = d2 * iv;
        e1[e] = r1;
        e2[e] = r2;
        lastinv = iv;
    }
    {
        double previnv = shup<W>(lastinv, 1);
        if (lane == 0) previnv = 0.0;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            am[e] *= previnv;
            previnv = inv[e];
        }
    }
    // ---- 2. forward elimination (affine scan)
    double A = 1.0;
    double2 B = make_double2(0.0, 0.0);
#pragma unroll
    for (int e = 0; e < E; ++e) {
        A = am[e] * A;
        B = make_double2(am[e] * B.x + f[e].x, am[e] * B.y + f[e].y);
    }
#pragma unroll
    for (int k = 1; k < W; k <<= 1) {
        const double oA = shup<W>(A, k), oBx = shup<W>(B.x, k), oBy = shup<W>(B.y, k);
        if (lane >= k) {
            B = make_double2(A * oBx + B.x, A * oBy + B.y);
            A = A * oA;
        }
    }
    double2 y = make_double2(shup<W>(B.x, 1), shup<W>(B.y, 1));
    if (lane == 0) y = make_double2(0.0, 0.0);
    double2 yp[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        y = make_double2(f[e].x + am[e] * y.x, f[e].y + am[e] * y.y);
        yp[e] = make_double2(y.x * inv[e], y.y * inv[e]);
    }
    // ---- 3. back substitution (2x2 affine suffix scan)
    // Phi = [[-(e1-e2), -e2], [-(e1-e2), 1-e2]], b = (yp, yp)
    double P00 = 1.0, P01 = 0.0, P10 = 0.0, P11 = 1.0;
    double2 c0 = make_double2(0.0, 0.0), c1 = make_double2(0.0, 0.0);
#pragma unroll
    for (int e = E - 1; e >= 0; --e) {
        const double q = -(e1[e] - e2[e]), w = -e2[e], w1 = 1.0 - e2[e];
        const double n00 = q * P00 + w * P10, n01 = q * P01 + w * P11;
        const double n10 = q * P00 + w1 * P10, n11 = q * P01 + w1 * P11;
        const double2 nc0 = make_double2(q * c0.x + w * c1.x + yp[e].x, q * c0.y + w * c1.y + yp[e].y);
        const double2 nc1 = make_double2(q * c0.x + w1 * c1.x + yp[e].x, q * c0.y + w1 * c1.y + yp[e].y);
        P00 = n00; P01 = n01; P10 = n10; P11 = n11;
        c0 = nc0;
        c1 = nc1;
    }
#pragma unroll
    for (int k = 1; k < W; k <<= 1) {
        const double o00 = shdn<W>(P00, k), o01 = shdn<W>(P01, k), o10 = shdn<W>(P10, k), o11 = shdn<W>(P11, k);
        const double oc0x = shdn<W>(c0.x, k), oc0y = shdn<W>(c0.y, k), oc1x = shdn<W>(c1.x, k),
                     oc1y = shdn<W>(c1.y, k);
        if (lane + k < W) {
            // (P, c) o (O, oc) = (P O, P oc + c)
            const double2 nc0 = make_double2(P00 * oc0x + P01 * oc1x + c0.x, P00 * oc0y + P01 * oc1y + c0.y);
            const double2 nc1 = make_double2(P10 * oc0x + P11 * oc1x + c1.x, P10 * oc0y + P11 * oc1y + c1.y);
            const double n00 = P00 * o00 + P01 * o10, n01 = P00 * o01 + P01 * o11;
            const double n10 = P10 * o00 + P11 * o10, n11 = P10 * o01 + P11 * o11;
            P00 = n00; P01 = n01; P10 = n10; P11 = n11;
            c0 = nc0;
            c1 = nc1;
        }
    }
    double2 sv = make_double2(shdn<W>(c0.x, 1), shdn<W>(c0.y, 1));
    double2 sT = make_double2(shdn<W>(c1.x, 1), shdn<W>(c1.y, 1));
    if (lane == W - 1) {
        sv = make_double2(0.0, 0.0);
        sT = make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int e = E - 1; e >= 0; --e) {
        const double q = -(e1[e] - e2[e]), w = -e2[e], w1 = 1.0 - e2[e];
        const double2 v = make_double2(q * sv.x + w * sT.x + yp[e].x, q * sv.y + w * sT.y + yp[e].y);
        const double2 T = make_double2(q * sv.x + w1 * sT.x + yp[e].x, q * sv.y + w1 * sT.y + yp[e].y);
        sv = v;
        sT = T;
        const int i = 2 * (m0 + e) + p;
        if (i < M) {
            double2 out = v;
            if (nm) {
                if (i == 0) {
                    // gauge of la.pyx:907-909 on the first LU pivot d0[0]
                    const double d00 = 1.0 / inv[0];
                    if ((d00 - 1.0) * (d00 - 1.0) < 1e-16) out = make_double2(0.0, 0.0);
                } else {
                    const double s = 1.0 / ((double)i * i);
                    out = make_double2(out.x * s, out.y * s);
                }
            }
            X[i] = out;
        }
    }
}

// Variant reading the right-hand side from X (rows i = 2m + p) first.
template <int E, int W = 16>
__device__ inline void helm_chain_scan(double2* X, int p, int M, double al, double be, bool nm,
                                       const HelmTab& tab, int lane)
{
    double2 f[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int i = 2 * (lane * E + e) + p;
        f[e] = (i < M) ? X[i] : make_double2(0.0, 0.0);
    }
    helm_chain_scan_f<E, W>(f, X, p, M, al, be, nm, tab, lane);
}

}  // namespace sfa

namespace sfa {

// Row i of (alfa A + beta B) for alfa = -1 from the closed-form Chebyshev
// SD/SN Galerkin diagonals (chebyshev/matrices.py:221-281, 377-427,
// 1047-1165), with the Neumann 1/j^2 column scaling and k = 0 gauge row of
// LU_Helmholtz_1D (la.pyx:735-749); the same values as helm_row on the
// uploaded tables, computed in registers.  Rows i >= M are identity padding.
__device__ inline HelmRow helm_row_cf(int i, int M, double be, bool nm, bool zb)
{
    const double pi = 3.14159265358979323846, hp = 0.5 * pi;
    HelmRow r;
    if (i >= M) {
        r.h = 1.0;
        r.s = r.c = r.g = 0.0;
        return r;
    }
    const double k = (double)i;
    const double hn0 = (i == 0) ? pi : hp;
    double a0, a2 = 0.0, a4 = 0.0, b0, b2 = 0.0, bm = 0.0;
    if (!nm) {
        a0 = -2.0 * pi * (k + 1.0) * (k + 2.0);
        if (i < M - 2) {
            a2 = -4.0 * pi * (k + 1.0);
            b2 = -hp;
        }
        if (i < M - 4) a4 = -4.0 * pi * (k + 1.0);
        b0 = hn0 + hp;
        if (i >= 2) bm = -hp;
    } else {
        const double q = k / (k + 2.0), al = q * q;
        const double k2i = 1.0 / ((k + 2.0) * (k + 2.0));
        if (i == 0 && zb) {
            a0 = -1.0;
            b0 = 0.0;
        } else {
            a0 = -2.0 * pi * k * k * (k + 1.0) / (k + 2.0);
            b0 = hn0 + al * al * hp;
            if (i > 0) {
                const double ii = 1.0 / (k * k);
                a0 *= ii;
                b0 *= ii;
            }
        }
        if (i < M - 2) {
            a2 = -4.0 * pi * (k + 2.0) * (k + 2.0) * (k + 1.0) * k2i * k2i;
            b2 = -al * hp * k2i;
        }
        if (i < M - 4) a4 = -4.0 * pi * (k + 4.0) * (k + 4.0) * (k + 1.0) * k2i / ((k + 4.0) * (k + 4.0));
        if (i >= 2) {
            const double qm = (k - 2.0) / k;
            bm = -(qm * qm) * hp;
            if (i - 2 >= 1) bm *= 1.0 / ((k - 2.0) * (k - 2.0));
        }
    }
    r.h = -a0 + be * b0;
    r.s = -a2 + be * b2;
    r.c = -a4;
    r.g = be * bm;
    return r;
}

// Device view of a fused-solve plan: line -> entry map, s = ky^2 + kz^2 per
// entry, segment restart states per (entry, parity), and the operator's
// diagonals.
struct XPlan {
    const int* linemap;   // (ky * Nzh + kz) -> entry
    const double* svals;  // [nu]
    const double4* bs;    // [nu][2][16]
};

// One row of the LU recurrence of LU_Helmholtz_1D (la.pyx:751-763) in ratio
// form.  State in: (r1, r2, pinv) = (d1/d0, d2/d0, 1/d0) of the previous row of
// the chain (first = first row of the chain).  Out: this row's 1/d0, the new
// ratios, and the forward-elimination multiplier am = -L.
// 1/d for the LU rows: hardware reciprocal refined by two Newton steps
// (5 instructions against the IEEE division's scale/fixup sequence; within an
// ulp for the normal-range pivots here).  The plan builder and the x-pass both
// regenerate rows through helm_lu_row, so their states agree.
__device__ inline double helm_rcp(double d)
{
    double r = __builtin_amdgcn_rcp(d);
    double e = __builtin_fma(-d, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-d, r, 1.0);
    return __builtin_fma(r, e, r);
}

__device__ inline void helm_lu_row(const HelmRow& r, int i, int M, bool first, double& r1, double& r2,
                                   double& pinv, double& iv, double& am)
{
    double d0, d1, d2;
    if (first) {
        d0 = r.h;
        d1 = r.s;
        d2 = r.c;
        am = 0.0;
    } else {
        d0 = r.h - r.g * r1;
        d1 = (i < M - 2) ? r.s - r.g * r2 : 0.0;
        d2 = (i < M - 4) ? r.c - r.g * r2 : 0.0;
        am = -r.g * pinv;
    }
    iv = helm_rcp(d0);
    r1 = d1 * iv;
    r2 = d2 * iv;
    pinv = iv;
}

// Segment restart states of the LU recurrence.  The chain's K rows are cut
// into 16 segments of E rows (one per lane of the solve).  bs[(2u + p) 16 + l]
// = (d1/d0, d2/d0, 1/d0, 0) of row l E - 1 of chain p (zero for l = 0), for
// alfa = -1, beta = alpha + s[u]: the sequential recurrence helm_lu_row,
// sampled at the segment boundaries.  One thread per (entry, parity).
static __global__ void helm_bstate_kernel(int M, int E, int neumann,
                                          double alpha, const double* __restrict__ svals,
                                          int64_t nu, double4* __restrict__ bs)
{
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= 2 * nu) return;
    const int64_t u = gid >> 1;
    const int p = (int)(gid & 1);
    const double be = alpha + svals[u];
    const bool nm = neumann != 0;
    const bool zb = nm && fabs(be) < 1e-8;
    double4* out = bs + (2 * u + p) * 16;
    out[0] = make_double4(0.0, 0.0, 0.0, 0.0);
    double r1 = 0.0, r2 = 0.0, pinv = 0.0, iv, am;
    for (int m = 0; m < 16 * E; ++m) {
        if (m > 0 && m % E == 0) out[m / E] = make_double4(r1, r2, pinv, 0.0);
        const int i = 2 * m + p;
        const HelmRow r = helm_row_cf(i, M, be, nm, zb);
        helm_lu_row(r, i, M, m == 0, r1, r2, pinv, iv, am);
    }
}

// Solve one chain from the segment restart state bs = (d1/d0, d2/d0, 1/d0) of
// the row before this lane's segment: the lane regenerates its E rows of LU
// factors with the exact row recurrence, then forward elimination and back
// substitution as affine scans over the W lanes.  X: the line, any storage
// type C2, holding row i at i + i / PADW (one pad element every PADW rows;
// PADW = 0: none), which keeps the lanes' rows (stride 2E) in distinct LDS
// banks.
template <int E, int W, typename C2, int PADW = 0>
__device__ inline void helm_chain_solve_bs(const double2 (&f)[E], C2* X, int p, int M, double be,
                                           bool nm, double4 bs, int lane)
{
    const bool zb = nm && fabs(be) < 1e-8;
    const int m0 = lane * E;
    double r1 = bs.x, r2 = bs.y, pinv = bs.z;
    // Forward elimination y_m = f_m - (g_m / d0_{m-1}) y_{m-1} in the scaled
    // unknowns z_m = y_m / d0_m:  z_m = f_m / d0_m + am_m z_{m-1} with
    // am_m = -g_m / d0_m, so a row keeps f / d0, am, d1/d0 and d2/d0 and no
    // separate pivot for the back substitution's y / d0.
    double2 yp[E];
    double e1[E], e2[E], am[E];
    bool gauge = false;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int i = 2 * (m0 + e) + p;
        const HelmRow r = helm_row_cf(i, M, be, nm, zb);
        double iv, amr;
        helm_lu_row(r, i, M, i < 2, r1, r2, pinv, iv, amr);
        am[e] = (i < 2) ? 0.0 : -r.g * iv;
        yp[e] = make_double2(f[e].x * iv, f[e].y * iv);
        e1[e] = r1;
        e2[e] = r2;
        if (nm && i == 0) gauge = (r.h - 1.0) * (r.h - 1.0) < 1e-16;
    }
    // affine prefix scan over lanes
    double A = 1.0;
    double2 B = make_double2(0.0, 0.0);
#pragma unroll
    for (int e = 0; e < E; ++e) {
        A = am[e] * A;
        B = make_double2(am[e] * B.x + yp[e].x, am[e] * B.y + yp[e].y);
    }
#pragma unroll
    for (int k = 1; k < W; k <<= 1) {
        const double oA = shup<W>(A, k), oBx = shup<W>(B.x, k), oBy = shup<W>(B.y, k);
        if (lane >= k) {
            B = make_double2(A * oBx + B.x, A * oBy + B.y);
            A = A * oA;
        }
    }
    double2 y = make_double2(shup<W>(B.x, 1), shup<W>(B.y, 1));
    if (lane == 0) y = make_double2(0.0, 0.0);
#pragma unroll
    for (int e = 0; e < E; ++e) {
        y = make_double2(yp[e].x + am[e] * y.x, yp[e].y + am[e] * y.y);
        yp[e] = y;
    }
    double P00 = 1.0, P01 = 0.0, P10 = 0.0, P11 = 1.0;
    double2 c0 = make_double2(0.0, 0.0), c1 = make_double2(0.0, 0.0);
#pragma unroll
    for (int e = E - 1; e >= 0; --e) {
        const double q = -(e1[e] - e2[e]), w = -e2[e], w1 = 1.0 - e2[e];
        const double n00 = q * P00 + w * P10, n01 = q * P01 + w * P11;
        const double n10 = q * P00 + w1 * P10, n11 = q * P01 + w1 * P11;
        const double2 nc0 = make_double2(q * c0.x + w * c1.x + yp[e].x, q * c0.y + w * c1.y + yp[e].y);
        const double2 nc1 = make_double2(q * c0.x + w1 * c1.x + yp[e].x, q * c0.y + w1 * c1.y + yp[e].y);
        P00 = n00; P01 = n01; P10 = n10; P11 = n11;
        c0 = nc0;
        c1 = nc1;
    }
#pragma unroll
    for (int k = 1; k < W; k <<= 1) {
        const double o00 = shdn<W>(P00, k), o01 = shdn<W>(P01, k), o10 = shdn<W>(P10, k), o11 = shdn<W>(P11, k);
        const double oc0x = shdn<W>(c0.x, k), oc0y = shdn<W>(c0.y, k), oc1x = shdn<W>(c1.x, k),
                     oc1y = shdn<W>(c1.y, k);
        if (lane + k < W) {
            const double2 nc0 = make_double2(P00 * oc0x + P01 * oc1x + c0.x, P00 * oc0y + P01 * oc1y + c0.y);
            const double2 nc1 = make_double2(P10 * oc0x + P11 * oc1x + c1.x, P10 * oc0y + P11 * oc1y + c1.y);
            const double n00 = P00 * o00 + P01 * o10, n01 = P00 * o01 + P01 * o11;
            const double n10 = P10 * o00 + P11 * o10, n11 = P10 * o01 + P11 * o11;
            P00 = n00; P01 = n01; P10 = n10; P11 = n11;
            c0 = nc0;
            c1 = nc1;
        }
    }
    double2 sv = make_double2(shdn<W>(c0.x, 1), shdn<W>(c0.y, 1));
    double2 sT = make_double2(shdn<W>(c1.x, 1), shdn<W>(c1.y, 1));
    if (lane == W - 1) {
        sv = make_double2(0.0, 0.0);
        sT = make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int e = E - 1; e >= 0; --e) {
        const double q = -(e1[e] - e2[e]), w = -e2[e], w1 = 1.0 - e2[e];
        const double2 v = make_double2(q * sv.x + w * sT.x + yp[e].x, q * sv.y + w * sT.y + yp[e].y);
        const double2 T = make_double2(q * sv.x + w1 * sT.x + yp[e].x, q * sv.y + w1 * sT.y + yp[e].y);
        sv = v;
        sT = T;
        const int i = 2 * (m0 + e) + p;
        if (i < M) {
            double2 out = v;
            if (nm) {
                if (i == 0) {
                    if (gauge) out = make_double2(0.0, 0.0);
                } else {
                    const double s = 1.0 / ((double)i * i);
                    out = make_double2(out.x * s, out.y * s);
                }
            }
            X[PADW ? i + i / PADW : i] = from_c64<C2>(out);
        }
    }
}

}  // namespace sfa
