// Shared device code for the relaxing-TVD MHD kernels (gfx950).
//
// Numerics follow the reference C path (Heterogeneous_MHD/C) operation by
// operation: FP32 arithmetic with the FP64 sub-expressions the C code gets
// from double literals and fabs/sqrt, no FMA contraction (-ffp-contract=off),
// IEEE FP32 division and square root, FP32 denormals kept.  With T = double
// every operation is FP64 (the FP64 instantiation used as an error yardstick).
//
// The one liberty taken: the reference's helper a*X + b*Y + c
// (Heterogeneous_MHD/C/includefunc_spu.h:73-94) is always called with
// c = 0; that "+ 0" only turns an exact -0 into +0, so it is left out.
// Results therefore equal the reference's numerically (-0 == +0, NaN where
// the reference has NaN); every non-zero value matches bit for bit.  (The
// "0*w" terms of its calls with b = 0 are kept: they make an infinite w
// give NaN.)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mhd {

// Pencil frame of a sweep over an x-major AoS box of nx*ny*nz cells.
// The sweep runs along i; (j,k) number the pencils.  Frame 0 is the x sweep
// (i,j,k) = (x,y,z); frame 1 the y sweep (i,j,k) = (y,x,z) with the reference's
// transpose12 component relabelling (u: 1,3,2,4,5, b: 2,1,3); frame 2 the z
// sweep (i,j,k) = (z,x,y), the state the reference reaches after transpose12
// then transpose13 (u: 1,4,2,3,5, b: 3,1,2).
struct Frame {
    int ni, nj, nk;     // pencil length, pencils along j, pencils along k
    long si, sj, sk;    // strides of i, j, k in cells
};

// SoA layout of one box: cell (i,j,k) at i + j*sy + k*sz of each plane,
// planes N apart; sy >= nx and sz >= sy*ny leave padding that breaks
// power-of-two strides
struct SoaGrid {
    int nx, ny, nz;
    long sy, sz, N;
};

// Frame component maps: frame component c lives in physical component map[c].
template <int FR> struct Comp;
template <> struct Comp<0> {
    static constexpr int u[5] = {0, 1, 2, 3, 4};
    static constexpr int b[3] = {0, 1, 2};
};
template <> struct Comp<1> {
    static constexpr int u[5] = {0, 2, 1, 3, 4};
    static constexpr int b[3] = {1, 0, 2};
};
template <> struct Comp<2> {
    static constexpr int u[5] = {0, 3, 1, 2, 4};
    static constexpr int b[3] = {2, 0, 1};
};
// inverse maps: physical component p is frame component inv[p]
template <int FR> struct CompInv;
template <> struct CompInv<0> {
    static constexpr int u[5] = {0, 1, 2, 3, 4};
    static constexpr int b[3] = {0, 1, 2};
};
template <> struct CompInv<1> {
    static constexpr int u[5] = {0, 2, 1, 3, 4};
    static constexpr int b[3] = {1, 0, 2};
};
template <> struct CompInv<2> {
    static constexpr int u[5] = {0, 2, 3, 1, 4};
    static constexpr int b[3] = {1, 2, 0};
};

// physical plane of frame component c, as selects (c may vary per lane)
template <int FR> __device__ __forceinline__ int cu(int c)
{
    return c == 0 ? Comp<FR>::u[0]
                  : c == 1 ? Comp<FR>::u[1] : c == 2 ? Comp<FR>::u[2] : c == 3 ? Comp<FR>::u[3] : Comp<FR>::u[4];
}
template <int FR> __device__ __forceinline__ int cb(int c)
{
    return c == 0 ? Comp<FR>::b[0] : c == 1 ? Comp<FR>::b[1] : Comp<FR>::b[2];
}

template <typename T> __device__ __forceinline__ T gamma_c() { return (T)(5.0 / 3.0); }

// The FP32 limiter's quotient (float)(N / D), N = 2.0*(double)p*(double)m
// (exact: a 48-bit product), D = (double)(p + m) with the sum rounded in
// FP32, without the operand scaling and special-case fixup of the IEEE
// double division:
//  * v_rcp_f64 of any float-valued D is within 2^-24.3 of 1/D (measured
//    over every finite nonzero float on MI355X; mhd_selfcheck repeats it);
//    one Newton step brings y within 2^-48 of 1/D;
//  * q = N*y, the remainder r = N - D*q is exact (fma), and q + r*y is
//    within 2^-96 of N/D, so the rounded q' lies within 2^-53 of N/D and
//    equals it whenever N/D is a double;
//  * N/D is either a float rounding midpoint (25 bits: then q' is exactly
//    it) or more than 2^-49 away from every midpoint in relative terms
//    (N has 48 significant bits, D 24), so q' and the IEEE quotient round
//    to the same float.
// The ranges involved (|N| in [2^-297, 2^257], |D| in [2^-149, 2^128])
// keep every step clear of double overflow and underflow.
__device__ __forceinline__ double limiter_refined(double N, double D)
{
    double y = __builtin_amdgcn_rcp(D);
    const double e = fma(-D, y, 1.0);
    y = fma(y, e, y);
    const double q = N * y;
    const double r = fma(-D, q, N);
    return fma(r, y, q);
}
// N = 2.0*(double)p*(double)m for FP32 p, m: exact (a 48-bit product).
__device__ __forceinline__ double limiter_numerator(float p, float m) { return 2.0 * (double)p * (double)m; }
// The limiter's test p*m > 0 (in FP32) from N: the FP32 product rounds to a
// positive value exactly when p*m > 2^-150 (2^-150 itself is the tie
// between 0 and the smallest subnormal, which goes to 0), infinities and
// FP32 overflow included, and NaN fails both.  One compare on N, which the
// quotient needs anyway, instead of an FP32 product and a compare.
__device__ __forceinline__ bool limiter_taken(double N) { return N > 0x1p-149; }
// An infinite D (p + m overflowed) takes the IEEE division.
__device__ __forceinline__ float limiter_quotient(float p, float m, double N)
{
    const float s = p + m;
    const double D = (double)s;
    double q;
    if (__builtin_isfinite(s))
        q = limiter_refined(N, D);
    else
        q = N / D;
    return (float)q;
}
// The same without a branch, for p*m > 0: an infinite sum (p + m
// overflowed in FP32, or p or m infinite) enters as D = +-2^1000.  For
// finite N the quotient is then below 2^-740 and rounds to the float zero
// of the sign of D, which is N/(+-inf); for N = +inf (an infinite operand)
// q is infinite, the remainder NaN, and so is the result, as inf/inf.
__device__ __forceinline__ float limiter_quotient_nb(float p, float m, double N)
{
    const float s = p + m;
    return (float)limiter_refined(N, fmin(fmax((double)s, -0x1p1000), 0x1p1000));
}

// IEEE FP64 square root for x = 0, +inf or x >= 2^-767: the compiler's own
// sequence (reciprocal square root estimate, a coupled Newton step on the
// root and the half reciprocal, two residual corrections; zero and +inf pass
// through) without its rescaling of arguments below 2^-767.  The CFL speed
// takes roots of FP64 quotients of floats, never that small.  Both
// sequences round correctly, so they agree on every argument.
__device__ __forceinline__ double sqrt_ranged(double x)
{
    const double r = __builtin_amdgcn_rsq(x);
    double g = x * r, h = r * 0.5;
    const double e = fma(-h, g, 0.5);
    g = fma(g, e, g);
    h = fma(h, e, h);
    double d = fma(-g, g, x);
    g = fma(d, h, g);
    d = fma(-g, g, x);
    g = fma(d, h, g);
    return (x == 0.0 || x == __builtin_inf()) ? x : g;
}
template <typename T> __device__ __forceinline__ double sqrt_of(double x)
{
    if constexpr (sizeof(T) == 4)
        return sqrt_ranged(x);
    else
        return sqrt(x);
}
// The same for a float-valued x (the freezing speed's root of an FP32
// quotient, mhdflux.c:45): the coupled Newton step on the root and one
// residual correction, without the half reciprocal's refinement and the
// second correction.  For every positive finite float x this is the
// correctly rounded double root (exhaustive scan on MI355X, 0 mismatches;
// mhd_selfcheck_sqrt repeats it); zero and +inf pass through.
__device__ __forceinline__ double sqrt_float_valued(double x)
{
    const double r = __builtin_amdgcn_rsq(x);
    double g = x * r;
    const double h = r * 0.5;
    const double e = fma(-h, g, 0.5);
    g = fma(g, e, g);
    const double d = fma(-g, g, x);
    g = fma(d, h, g);
    return (x == 0.0 || x == __builtin_inf()) ? x : g;
}
template <typename T> __device__ __forceinline__ double sqrt_of_fv(double x)
{
    if constexpr (sizeof(T) == 4)
        return sqrt_float_valued(x);
    else
        return sqrt(x);
}

// van Leer harmonic mean: the test is the product p*m > 0 in T (for FP32
// read off the exact FP64 numerator, limiter_taken), the value 2*p*m/(p+m)
// in double with p+m summed in T (tvd1.c:71-81, tvdb.c:72-76).
template <typename T> __device__ __forceinline__ T vanleer(T p, T m)
{
    T d = (T)0;
    if constexpr (sizeof(T) == 4) {
        const double N = limiter_numerator(p, m);
        if (limiter_taken(N)) d = limiter_quotient(p, m, N);
    } else {
        if (p * m > (T)0) d = (T)(2.0 * (double)p * (double)m / (double)(p + m));
    }
    return d;
}
// The same with the FP32 quotient formed on every lane and selected where
// p*m > 0: no branch (the streamed tvdb, whose waves run it once per row)
template <typename T> __device__ __forceinline__ T vanleer_nb(T p, T m)
{
    if constexpr (sizeof(T) == 4) {
        const double N = limiter_numerator(p, m);
        const float d = limiter_quotient_nb(p, m, N);
        return limiter_taken(N) ? d : 0.0f;
    } else {
        return vanleer<T>(p, m);
    }
}

// Exact shortcuts of three FP64 sub-expressions of the reference's FP32 code
// (double literals promote them).  For FP32 operands:
//  * (float)((double)x / 2.0): the halving is exact in double, so this is
//    x/2 rounded once to float, which the FP32 product x*0.5f also is
//    (FP32 denormals are kept): one operation instead of three;
//  * (float)((double)a - (double)x / 2.0): x/2 is exact, so this is the
//    difference of two numbers of 24 significant bits rounded to double and
//    then to float.  That double rounding is innocuous: where a - x/2 is not
//    a double, the exponents differ by 29 or more and the result is within
//    1/32 of an ulp of the larger operand, which both roundings return.  So
//    it is a - x/2 rounded once to float, the FP32 fma(x, -0.5, a): one
//    operation instead of four;
//  * (float)(((double)s + 2.0*(double)v) / 4.0): 2v is exact, so the sum
//    is one fma, and /4 is exact.
// The FP64 instantiation keeps the literal expressions.
template <typename T> __device__ __forceinline__ T half_of(T x)
{
    if constexpr (sizeof(T) == 4)
        return x * 0.5f;
    else
        return (T)((double)x / 2.0);
}
template <typename T> __device__ __forceinline__ T sub_half(T a, T x)
{
    if constexpr (sizeof(T) == 4)
        return fmaf(x, -0.5f, a);
    else
        return (T)((double)a - (double)x / 2.0);
}
template <typename T> __device__ __forceinline__ T avg121(T s, T v)
{
    if constexpr (sizeof(T) == 4)
        return (T)(fma((double)v, 2.0, (double)s) * 0.25);
    else
        return (T)(((double)s + 2.0 * (double)v) / 4.0);
}

// ps of mhdflux.c:28 / calcfl.c:50 from q = S/u1 (S = u2^2+u3^2+u4^2) and
// B = b1^2+b2^2+b3^2 in T:
//   ((double)u5 - (double)q/2.0)*((double)g - 1.0) + (2.0 - (double)g)*(double)B/2.0
// For FP32 two steps take one operation each instead of two, with the same
// result: (double)q/2.0 is exact (a float value halves exactly in double),
// so u5 - q/2 is one fma; and (2 - g) has 24 significant bits, so
// (2 - g)*B is exact in double and halving it is exact, which makes it
// B*((2 - g)/2), one product.
template <typename T> __device__ __forceinline__ T pstar_q(T q, T u5, T B)
{
    const T g = gamma_c<T>();
    if constexpr (sizeof(T) == 4) {
        const double t = fma((double)q, -0.5, (double)u5);
        return (T)(t * ((double)g - 1.0) + (double)B * ((2.0 - (double)g) * 0.5));
    } else {
        return (T)(((double)u5 - (double)q / 2.0) * ((double)g - 1.0) + (2.0 - (double)g) * (double)B / 2.0);
    }
}
template <typename T> __device__ __forceinline__ T pstar(T u1, T u5, T S, T B) { return pstar_q<T>(S / u1, u5, B); }
// p = ps - B/2.0 in double (mhdflux.c:37, calcfl.c:52); for FP32 this is
// ps - B/2 rounded once to float, as sub_half: one FP32 fma
template <typename T> __device__ __forceinline__ T pgas(T ps, T B)
{
    if constexpr (sizeof(T) == 4)
        return fmaf(B, -0.5f, ps);
    else
        return (T)((double)ps - (double)B / 2.0);
}

// One cell of mhdflux (mhdflux.c:21-46): unnormalised flux v[5] and the
// freezing speed ci.
template <typename T>
__device__ __forceinline__ void mhdflux_cell(const T *u, const T *b, T *v, T &ci)
{
    const T u1 = u[0], u2 = u[1], u3 = u[2], u4 = u[3], u5 = u[4];
    const T b1 = b[0], b2 = b[1], b3 = b[2];
    const T vx = u2 / u1;
    const T S = u2 * u2 + u3 * u3 + u4 * u4;
    const T B = b1 * b1 + b2 * b2 + b3 * b3;
    const T ps = pstar<T>(u1, u5, S, B);
    v[0] = u2;
    v[1] = u2 * vx + ps - b1 * b1;
    v[2] = u3 * vx - b2 * b1;
    v[3] = u4 * vx - b3 * b1;
    v[4] = (u5 + ps) * vx - b1 * (b1 * u2 + b2 * u3 + b3 * u4) / u1;
    const T p = pgas<T>(ps, B);
    const T g = gamma_c<T>();
    ci = (T)(fabs((double)vx) + sqrt_of_fv<T>(fabs((double)((B + g * p) / u1))));
}

// ---- exact fast FP32 division ------------------------------------------
// a/d for several a sharing one divisor d: y = rcp(d) refined by one Newton
// step (y = fma(fma(-d, y0, 1), y0, y0)), then q = a*y and ONE residual
// correction q + fma(-d, q, a)*y.  The compiler's IEEE division runs the
// same reciprocal and quotient with a second correction (v_div_scale ...
// v_div_fmas); the first already gives the correctly rounded quotient:
//  * for a, d in [1, 2) this holds for every one of the 2^46 mantissa pairs
//    (exhaustive scan on MI355X, mhd_selfcheck_division);
//  * v_rcp_f32 is exactly odd and exactly scales by powers of two over
//    divisor exponents [-40, 40] (same scan), and the remaining operations
//    are IEEE mul/fma whose results stay normal for d of exponent in
//    [-31, 29] and a of exponent in [-61, 59]: y and q are within 2^90 of
//    1, and the residual a - d*q is a multiple of 2^(e_a - 47), so at
//    least 2^-108 when nonzero.  Every such pair is therefore a scanned
//    pair scaled by powers of two, with the same rounding.
// A zero a gives a zero quotient that may differ from a/d in sign only.
// Callers check those ranges in bulk and fall back to true divisions
// outside them.
struct Recip {
    float d, y;
};
__device__ __forceinline__ Recip recip_of(float d)
{
    const float y0 = __builtin_amdgcn_rcpf(d);
    return {d, fmaf(fmaf(-d, y0, 1.0f), y0, y0)};
}

// a/d from the refined reciprocal of d (ranges checked by the caller)
__device__ __forceinline__ float quot_by(float a, const Recip &r)
{
    const float q = a * r.y;
    return fmaf(fmaf(-r.d, q, a), r.y, q);
}

// mhdflux_cell (FP32) with the four divisions by u1 through one refined
// reciprocal (quot_by); the operand ranges are not checked here but folded
// into the caller's bulk check: ok takes |u1| in [2^-31, 2^30), vmax the
// largest |numerator| and emin the smallest numerator exponent.
__device__ __forceinline__ void mhdflux_cell_bulk(const float *u, const float *b, float *v, float &ci, bool &ok,
                                                  float &vmax, int &emin)
{
    const float u1 = u[0], u2 = u[1], u3 = u[2], u4 = u[3], u5 = u[4];
    const float b1 = b[0], b2 = b[1], b3 = b[2];
    ok = ok && fabsf(u1) < 0x1p30f && __builtin_amdgcn_frexp_expf(u1) >= -30;
    const Recip r1 = recip_of(u1);
    const float vx = quot_by(u2, r1);
    const float S = u2 * u2 + u3 * u3 + u4 * u4;
    const float B = b1 * b1 + b2 * b2 + b3 * b3;
    const float g = gamma_c<float>();
    const float ps = pstar_q<float>(quot_by(S, r1), u5, B);
    v[0] = u2;
    v[1] = u2 * vx + ps - b1 * b1;
    v[2] = u3 * vx - b2 * b1;
    v[3] = u4 * vx - b3 * b1;
    const float bm = b1 * (b1 * u2 + b2 * u3 + b3 * u4);
    v[4] = (u5 + ps) * vx - quot_by(bm, r1);
    const float p = pgas<float>(ps, B);
    const float w = B + g * p;
    ci = (float)(fabs((double)vx) + sqrt_float_valued(fabs((double)quot_by(w, r1))));
    vmax = fmaxf(fmaxf(vmax, fabsf(u2)), fmaxf(fmaxf(fabsf(S), fabsf(bm)), fabsf(w)));
    emin = min(min(emin, min(__builtin_amdgcn_frexp_expf(u2), __builtin_amdgcn_frexp_expf(S))),
               min(__builtin_amdgcn_frexp_expf(bm), __builtin_amdgcn_frexp_expf(w)));
}

// Per-cell fast speed of calcfl.c:45-54 (B^2 of the face values at the cell,
// or of the cell averages calcfl.c:38-43 forms when the caller passes them).
template <typename T>
__device__ __forceinline__ T cfl_speed(T u1, T u2, T u3, T u4, T u5, T B)
{
    const T g = gamma_c<T>();
    // calcfl.c:45-49 takes the largest of |u2/u1|, |u3/u1|, |u4/u1| by the
    // comparisons of max_3num (:67-91).  For a finite nonzero u1 a correctly
    // rounded division is monotonic in the dividend, so that is the quotient
    // of the largest |u_k| picked by the same comparisons (two dividends
    // that compare unequal may give equal quotients, but then either pick
    // has that value; a NaN dividend gives a NaN quotient and loses every
    // comparison either way): one division instead of three.  For u1 = 0 or
    // infinite the quotients 0/0 and inf/inf are NaN while their dividends
    // are not, so the three quotients are compared as the reference does.
    const T a1 = fabs(u2), a2 = fabs(u3), a3 = fabs(u4);
    T v;
    if (__builtin_isfinite(u1) && u1 != (T)0) {
        v = (a1 > a2 ? (a1 > a3 ? a1 : a3) : (a2 > a3 ? a2 : a3)) / fabs(u1);
    } else {
        const T q1 = fabs(u2 / u1), q2 = fabs(u3 / u1), q3 = fabs(u4 / u1);
        v = q1 > q2 ? (q1 > q3 ? q1 : q3) : (q2 > q3 ? q2 : q3);
    }
    const T S = u2 * u2 + u3 * u3 + u4 * u4;
    const T ps = pstar<T>(u1, u5, S, B);
    const T p = pgas<T>(ps, B);
    // B*2.0 + gamma*p in double; for FP32 2B is exact, so the sum is one fma
    const double num = sizeof(T) == 4 ? fma((double)B, 2.0, (double)(g * p)) : (double)B * 2.0 + (double)(g * p);
    return (T)((double)v + sqrt_of<T>(fabs(num) / (double)u1));
}

// ---- wave-level helpers: a pencil of n = L*K cells lives in one wave, lane
// l < L owning cells l*K .. l*K+K-1 in registers. ----

__device__ __forceinline__ int lane_id() { return __lane_id(); }

template <typename T> __device__ __forceinline__ T shfl_from(T x, int src)
{
    if constexpr (sizeof(T) == 4) {
        return __int_as_float(__builtin_amdgcn_ds_bpermute(src << 2, __float_as_int(x)));
    } else {
        const long long v = __double_as_longlong(x);
        const int lo = __builtin_amdgcn_ds_bpermute(src << 2, (int)(v & 0xffffffff));
        const int hi = __builtin_amdgcn_ds_bpermute(src << 2, (int)(v >> 32));
        return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
    }
}

// a[q] <- a[q+1] along the periodic pencil (cshift by +1)
template <typename T, int K>
__device__ __forceinline__ void shift_next(const T (&a)[K], T (&o)[K], int src_next)
{
#pragma unroll
    for (int q = 0; q < K - 1; q++) o[q] = a[q + 1];
    o[K - 1] = shfl_from(a[0], src_next);
}

// a[q] <- a[q-1] along the periodic pencil (cshift by -1)
template <typename T, int K>
__device__ __forceinline__ void shift_prev(const T (&a)[K], T (&o)[K], int src_prev)
{
    o[0] = shfl_from(a[K - 1], src_prev);
#pragma unroll
    for (int q = 1; q < K; q++) o[q] = a[q - 1];
}

// ---- lane geometry of a pencil ----
// Whole pencil: n = L*K cells on L lanes; next/prev are the neighbour lanes
// (periodic over L).
struct LanesFull {
    int next, prev;
    __device__ __forceinline__ bool valid(int) const { return true; }
};
// Ragged pencil: L = ceil(n/K) lanes, the last holding r = n - (L-1)*K < K
// cells.  wq is the lane's last cell (K-1, or r-1 in the last lane); cells
// beyond it hold values nothing valid reads (the shifts below wrap the
// pencil at wq), and they are kept out of every pencil-wide reduction.
struct LanesRagged {
    int next, prev, wq;
    __device__ __forceinline__ bool valid(int q) const { return q <= wq; }
};
__device__ __forceinline__ LanesRagged ragged_lanes(int n, int K, int lc)
{
    const int L = (n + K - 1) / K;
    LanesRagged g;
    g.next = lc + 1 == L ? 0 : lc + 1;
    g.prev = lc == 0 ? L - 1 : lc - 1;
    g.wq = lc == L - 1 ? n - (L - 1) * K - 1 : K - 1;
    return g;
}

// Several pencils per wave: 64/G pencils of n = G*K cells, pencil g on the
// aligned lanes g*G .. g*G+G-1 (lane l holds cells (l % G)*K .. of pencil
// l / G).  next/prev wrap inside the lane group, and the pencil max is a
// reduction over the group only.
template <int G> struct LanesGroup {
    static_assert(G >= 2 && G <= 32 && (G & (G - 1)) == 0, "a lane group is a power of two below the wave");
    int next, prev;
    __device__ __forceinline__ bool valid(int) const { return true; }
};
template <int G> __device__ __forceinline__ LanesGroup<G> group_lanes(int l)
{
    const int base = l & ~(G - 1), lc = l & (G - 1);
    return LanesGroup<G>{base + ((lc + 1) & (G - 1)), base + ((lc + G - 1) & (G - 1))};
}

template <typename T, int K> __device__ __forceinline__ void shift_next(const T (&a)[K], T (&o)[K], const LanesFull &g)
{
    shift_next<T, K>(a, o, g.next);
}
template <typename T, int K> __device__ __forceinline__ void shift_prev(const T (&a)[K], T (&o)[K], const LanesFull &g)
{
    shift_prev<T, K>(a, o, g.prev);
}
template <typename T, int K, int G>
__device__ __forceinline__ void shift_next(const T (&a)[K], T (&o)[K], const LanesGroup<G> &g)
{
    shift_next<T, K>(a, o, g.next);
}
template <typename T, int K, int G>
__device__ __forceinline__ void shift_prev(const T (&a)[K], T (&o)[K], const LanesGroup<G> &g)
{
    shift_prev<T, K>(a, o, g.prev);
}
// a[q] <- a[q+1]: the next lane's first cell lands at wq (lane 0's first cell
// in the last lane, which is where the pencil wraps)
template <typename T, int K>
__device__ __forceinline__ void shift_next(const T (&a)[K], T (&o)[K], const LanesRagged &g)
{
    const T x = shfl_from(a[0], g.next);
#pragma unroll
    for (int q = 0; q < K - 1; q++) o[q] = q == g.wq ? x : a[q + 1];
    o[K - 1] = x;
}
// a[q] <- a[q-1]: each lane sends its cell wq, so lane 0 receives the
// pencil's last cell
template <typename T, int K>
__device__ __forceinline__ void shift_prev(const T (&a)[K], T (&o)[K], const LanesRagged &g)
{
    T s = a[K - 1];
#pragma unroll
    for (int q = 0; q < K - 1; q++) s = q == g.wq ? a[q] : s;
    o[0] = shfl_from(s, g.prev);
#pragma unroll
    for (int q = 1; q < K; q++) o[q] = a[q - 1];
}

// Max over the 64 lanes (callers pass 0 from inactive lanes; all values >= 0).
// fmax skips NaN, like find_max's strict > (mhdflux.c:66-69).
template <typename T> __device__ __forceinline__ T wave_max(T x)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) x = fmax(x, shfl_from(x, lane_id() ^ off));
    return x;
}
// the pencil max of each lane geometry: the whole wave, or the lane group
template <typename T> __device__ __forceinline__ T pencil_max(T x, const LanesFull &) { return wave_max(x); }
template <typename T> __device__ __forceinline__ T pencil_max(T x, const LanesRagged &) { return wave_max(x); }
template <typename T, int G> __device__ __forceinline__ T pencil_max(T x, const LanesGroup<G> &)
{
#pragma unroll
    for (int off = G / 2; off >= 1; off >>= 1) x = fmax(x, shfl_from(x, lane_id() ^ off));
    return x;
}

// ---- the CFL max (calcfl.c:55, max_2num in k,j,i order) ----
// Every CFL speed is +0, positive, +inf or NaN (never -0: v >= +0 and the
// root is +0 or -0, whose sum with v is v or +0).  The reference's running
// max_2num(c, t) = c > t ? c : t resets to NaN at a NaN speed and to the
// next speed after it, so its result depends on where the NaN cells are.
// The parallel reductions therefore carry a NaN-sticky max - as unsigned
// bit patterns every NaN lies above +inf and the other speeds keep their
// order - and a NaN result sends the consumer to the exact sequential form
// (cfl_after_last_nan).  Without NaN both give the plain maximum.
template <typename T> __device__ __forceinline__ unsigned long long speed_bits(T x)
{
    if constexpr (sizeof(T) == 4)
        return (unsigned long long)__float_as_uint(x);
    else
        return (unsigned long long)__double_as_longlong(x);
}
template <typename T> __device__ __forceinline__ T speed_from_bits(unsigned long long x)
{
    if constexpr (sizeof(T) == 4)
        return __uint_as_float((unsigned)x);
    else
        return __longlong_as_double((long long)x);
}
template <typename T> __device__ __forceinline__ T cfl_max(T a, T b)
{
    if constexpr (sizeof(T) == 4)
        return __uint_as_float(max(__float_as_uint(a), __float_as_uint(b)));
    else
        return __longlong_as_double((long long)max((unsigned long long)__double_as_longlong(a),
                                                   (unsigned long long)__double_as_longlong(b)));
}
template <typename T> __device__ __forceinline__ T wave_cfl_max(T x)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) x = cfl_max(x, shfl_from(x, lane_id() ^ off));
    return x;
}

// The reference's sequential CFL max over ncell speeds (speed(s) for the
// cell of linear index s = (k*ny+j)*nx+i) when some speed is NaN: the max
// of the speeds after the last NaN one, floored at the initial c = 0 (all
// speeds are >= +0), or NaN when the last cell's speed is NaN.  Called by
// every thread of the block; two passes over the box (the rare path of a
// state that already holds NaN or a non-positive density).
template <typename T, class Speed> __device__ T cfl_after_last_nan(long ncell, const Speed &speed)
{
    __shared__ unsigned long long last1, mbits;  // last NaN index + 1 (0: none); max speed bits
    if (threadIdx.x == 0) {
        last1 = 0ull;
        mbits = 0ull;
    }
    __syncthreads();
    long mine = -1;
    for (long s = threadIdx.x; s < ncell; s += blockDim.x) {
        const T t = speed(s);
        if (t != t) mine = s;
    }
    if (mine >= 0) atomicMax(&last1, (unsigned long long)(mine + 1));
    __syncthreads();
    const long first = (long)last1;  // the cells after the last NaN one
    T c = (T)0;
    for (long s = first + threadIdx.x; s < ncell; s += blockDim.x) c = fmax(c, speed(s));
    atomicMax(&mbits, speed_bits(c));
    __syncthreads();
    const T r = first == ncell ? (T)__builtin_nan("") : speed_from_bits<T>(mbits);
    __syncthreads();  // the shared words are read before any later call resets them
    return r;
}

// The block's 6-double monitor record (sums of the first five values, max
// of the sixth) through LDS in a fixed order, without per-wave shuffles:
// every thread writes its values to part (6 * blockDim.x doubles, value-
// major), then thread (v, g) of the first 6 * G threads (G = blockDim.x / 16)
// folds values g, g + G, ..., g + 15 G of value v, and thread v folds those
// G partial results.  Deterministic: the order depends on the block shape only.
// The caller makes part free before (a barrier after its last other use);
// blockDim.x is a multiple of 64.
__device__ __forceinline__ void monitor_block_store_lds(const double (&ms)[6], double *dst, double *part)
{
    const int T = blockDim.x, t = threadIdx.x, G = T >> 4;
#pragma unroll
    for (int v = 0; v < 6; v++) part[v * T + t] = ms[v];
    __syncthreads();
    double r = 0.0;
    if (t < 6 * G) {
        const int v = t / G, g = t - v * G;
        const double *pv = part + v * T + g;
        r = pv[0];
        if (v < 5) {
#pragma unroll
            for (int i = 1; i < 16; i++) r = r + pv[i * G];
        } else {
#pragma unroll
            for (int i = 1; i < 16; i++) r = fmax(r, pv[i * G]);
        }
    }
    __syncthreads();
    if (t < 6 * G) part[t] = r;
    __syncthreads();
    if (t < 6) {
        const double *pv = part + t * G;
        double s = pv[0];
        for (int g = 1; g < G; g++) s = t < 5 ? s + pv[g] : fmax(s, pv[g]);
        dst[t] = s;
    }
}

// FP64 monitors of a block of cells: the lanes' partials ms[0..4] (mass,
// energy, momentum x, y, z) summed and ms[5] (max |div B|) max-ed, in a
// fixed order - an xor butterfly in each wave (both lanes of a pair add the
// same two values, so every lane ends with the same sum), then the waves in
// index order - so the result does not depend on timing.  Thread 0 writes
// dst[0..5].  Every thread of the block calls it.
__device__ __forceinline__ void monitor_block_store(const double (&ms)[6], double *dst)
{
    __shared__ double red[16][6];
    double a[6];
#pragma unroll
    for (int q = 0; q < 6; q++) a[q] = ms[q];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const int src = lane_id() ^ off;
#pragma unroll
        for (int q = 0; q < 5; q++) a[q] = a[q] + shfl_from(a[q], src);
        a[5] = fmax(a[5], shfl_from(a[5], src));
    }
    const int w = threadIdx.x >> 6;
    if (lane_id() == 0)
#pragma unroll
        for (int q = 0; q < 6; q++) red[w][q] = a[q];
    __syncthreads();
    if (threadIdx.x == 0) {
        double s[6];
#pragma unroll
        for (int q = 0; q < 6; q++) s[q] = red[0][q];
        for (int v = 1; v < (int)(blockDim.x >> 6); v++) {
#pragma unroll
            for (int q = 0; q < 5; q++) s[q] = s[q] + red[v][q];
            s[5] = fmax(s[5], red[v][5]);
        }
#pragma unroll
        for (int q = 0; q < 6; q++) dst[q] = s[q];
    }
}

// The CFL speed (calcfl.c:45-54) of the cell of linear index
// s = (k*ny+j)*nx+i, with the face B at the cell or (AVG) the cell average
// of calcfl.c:38-43: over a box in the step-level AoS layout, and over a
// device-resident SoA box
template <typename T, bool AVG> struct AosSpeed {
    const T *u, *b;
    int nx, ny, nz;
    __device__ T operator()(long s) const
    {
        const T *pu = u + 5 * s, *pb = b + 3 * s;
        T B;
        if constexpr (AVG) {
            const int i = (int)(s % nx);
            const long r = s / nx;
            const int j = (int)(r % ny), k = (int)(r / ny);
            const long sip = s - i + (i + 1 == nx ? 0 : i + 1);
            const long sjp = ((long)k * ny + (j + 1 == ny ? 0 : j + 1)) * nx + i;
            const long skp = ((long)(k + 1 == nz ? 0 : k + 1) * ny + j) * nx + i;
            const T bx = half_of<T>(pb[0] + b[3 * sip + 0]);
            const T by = half_of<T>(pb[1] + b[3 * sjp + 1]);
            const T bz = half_of<T>(pb[2] + b[3 * skp + 2]);
            B = bx * bx + by * by + bz * bz;
        } else {
            B = pb[0] * pb[0] + pb[1] * pb[1] + pb[2] * pb[2];
        }
        return cfl_speed<T>(pu[0], pu[1], pu[2], pu[3], pu[4], B);
    }
};
template <typename T, bool AVG> struct SoaSpeed {
    const T *u, *b;
    SoaGrid g;
    __device__ T operator()(long s) const
    {
        const int i = (int)(s % g.nx);
        const long r = s / g.nx;
        const int j = (int)(r % g.ny), k = (int)(r / g.ny);
        const long row = (long)j * g.sy + (long)k * g.sz, c = row + i, N = g.N;
        T B;
        if constexpr (AVG) {
            const long sip = row + (i + 1 == g.nx ? 0 : i + 1);
            const long sjp = (long)(j + 1 == g.ny ? 0 : j + 1) * g.sy + (long)k * g.sz + i;
            const long skp = (long)j * g.sy + (long)(k + 1 == g.nz ? 0 : k + 1) * g.sz + i;
            const T bx = half_of<T>(b[c] + b[sip]);
            const T by = half_of<T>(b[N + c] + b[N + sjp]);
            const T bz = half_of<T>(b[2 * N + c] + b[2 * N + skp]);
            B = bx * bx + by * by + bz * bz;
        } else {
            B = b[c] * b[c] + b[N + c] * b[N + c] + b[2 * N + c] * b[2 * N + c];
        }
        return cfl_speed<T>(u[c], u[N + c], u[2 * N + c], u[3 * N + c], u[4 * N + c], B);
    }
};

// Block index -> tile index so that each XCD (blocks are dealt round-robin
// over the 8 XCDs) receives a contiguous range of tiles: neighbouring tiles
// share cache lines and neighbour pencils in the same L2.
__device__ __forceinline__ int xcd_tile(int bid, int nb)
{
    const int q = nb >> 3, r = nb & 7;
    const int x = bid & 7, idx = bid >> 3;
    return x < r ? x * (q + 1) + idx : r * (q + 1) + (x - r) * q + idx;
}

}  // namespace mhd
