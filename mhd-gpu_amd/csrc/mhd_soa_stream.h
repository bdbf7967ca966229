// Column-streaming advectbyzx for the y and z frames of the SoA state.
//
// In those frames the pencils run along y or z and neighbouring pencils
// along j are neighbouring x, contiguous in memory.  Here a wave holds 64
// consecutive pencils, one per lane, and streams along the pencil row by
// row: every load and store of a row is one coalesced 256-byte access (the
// tile kernels read 16-byte rows there, one cache line per row, and were
// bound by the L1 request rate).  tvdb (Heterogeneous_MHD/C/tvdb.c:10-91)
// has no pencil-wide reduction, only a stencil of +-4 rows, so each lane
// carries a window of 15 values and emits one result per row, four rows
// behind the row it loads.
//
// The update is the reference's two halves (advectbyzx.c:14-23): pass A
// advects b2 along i with the face velocity between pencils j-1 and j
// (advectbyzxA1.c:34-56) and stores the new b2 and the EMF flux FA in a
// temporary plane (the reference's `temp`); pass B advects b3 with the face
// velocity between planes k-1 and k, takes the flux at the k+1 face from a
// second stream, and updates b1 = (((b1 - FA_j) + FA_{j+1}) - FB_k) + FB_{k+1}
// (A1.c:55, A2.c:23, B1.c:54, B2.c:23).  Every arithmetic expression is the
// tile kernels' (mhd_core.h tvdb_pencil), evaluated in the same order.
//
// A pencil is cut into segments of S rows for parallelism; each segment
// re-reads 4 rows on either side.
#pragma once
#include "mhd_soa.h"

namespace mhd {

template <typename T> __device__ __forceinline__ T face_vel(T ma, T mb, T ra, T rb) { return (ma + mb) / (ra + rb); }

// One tvdb pencil streamed row by row.  Row t brings the face-velocity
// operands (minus side a, plus side b: A1.c:34) and the field value at t;
// the push returns bnew[t-4] and F[t-4] = flux[t-5], the shifted flux of the
// CT update.  Valid from the ninth push on (the first eight fill the window).
template <typename T> struct TvdbStream {
    T vr2 = 0, vr1 = 0;                  // raw face velocity at t-2, t-1
    T vg2 = 0;                           // smoothed velocity at t-2
    T bw4 = 0, bw3 = 0, bw2 = 0, bw1 = 0;  // field at t-4 .. t-1
    bool up4 = false, up3 = false;       // vh > 0 at t-4, t-3 (vh is only ever tested for its sign)
    T f13 = 0;                           // half-step flux at t-3
    T g5 = 0, g4 = 0, g3 = 0;            // vg*b1 at t-5 .. t-3
    T fl5 = 0;                           // flux at t-5

    __device__ __forceinline__ void push(T ma, T mb, T ra, T rb, T bt, T dt, T &bnew, T &F)
    {
        const T vr0 = face_vel(ma, mb, ra, rb);
        // A1.c:41 vx = (vx[i-1] + vx[i+1] + 2.0*vx)/4.0 at i = t-1
        const T vg1 = avg121<T>(vr2 + vr0, vr1);
        // tvdb.c:27 vh = 0.5*vg + 0.5*vg[i+1] at i = t-2; tvdb.c:33 b*vg at t-1
        const bool up2 = (T)0.5 * vg2 + (T)0.5 * vg1 > (T)0;
        const T bv1 = bw1 * vg1;
        // tvdb.c:36-46 upwind half-step flux at t-2 (b*vg at t-2 formed
        // again from the stored field and velocity: the same product)
        const T bv2 = bw2 * vg2;
        const T f12 = up2 ? bv2 : bv1;
        // tvdb.c:51 b1 = b - (flux1 - flux1[i-1])*dt/2.0 at t-2
        const T b12 = sub_half<T>(bw2, (f12 - f13) * dt);
        const T g2 = vg2 * b12;
        // tvdb.c:54-78 limited flux at i = t-4 (g[i-1..i+2] = g5, g4, g3, g2)
        T w, wp, wm;
        if (up4) {
            w = g4;
            wp = half_of<T>(g3 - w);
            wm = half_of<T>(w - g5);
        } else {
            w = g3;
            wp = half_of<T>(w - g2);
            wm = half_of<T>(g4 - w);
        }
        const T dw = vanleer_nb<T>(wm, wp);
        const T fl4 = (w + dw) * dt;
        // tvdb.c:87 b = b - (flux - flux[i-1]) at t-4
        bnew = bw4 - (fl4 - fl5);
        F = fl5;
        shift(vr0, vg1, bt, up2, f12, g2);
        fl5 = fl4;
    }

    // A push that only fills the window: the first rows of a segment, whose
    // limited flux nothing reads (it takes the limiter from the eighth push
    // on, whose flux the first emitted row uses).  Leaves fl5 undefined.
    __device__ __forceinline__ void push_fill(T ma, T mb, T ra, T rb, T bt, T dt)
    {
        const T vr0 = face_vel(ma, mb, ra, rb);
        const T vg1 = avg121<T>(vr2 + vr0, vr1);
        const bool up2 = (T)0.5 * vg2 + (T)0.5 * vg1 > (T)0;
        const T bv1 = bw1 * vg1;
        const T bv2 = bw2 * vg2;
        const T f12 = up2 ? bv2 : bv1;
        const T b12 = sub_half<T>(bw2, (f12 - f13) * dt);
        const T g2 = vg2 * b12;
        shift(vr0, vg1, bt, up2, f12, g2);
    }

    __device__ __forceinline__ void shift(T vr0, T vg1, T bt, bool up2, T f12, T g2)
    {
        vr2 = vr1;
        vr1 = vr0;
        vg2 = vg1;
        bw4 = bw3;
        bw3 = bw2;
        bw2 = bw1;
        bw1 = bt;
        up4 = up3;
        up3 = up2;
        f13 = f12;
        g5 = g4;
        g4 = g3;
        g3 = g2;
    }
};

// periodic row index; t lies in [-4, n + 4) and n may be as small as 2
__device__ __forceinline__ int wrap_row(int t, int n)
{
    while (t < 0) t += n;
    while (t >= n) t -= n;
    return t;
}

// pass A: b2 along i with the j-face velocity; new b2 -> bout, FA -> fa
template <typename T, int FR>
__device__ __forceinline__ void advect_stream_a(const T *__restrict__ u, const T *__restrict__ bin,
                                                T *__restrict__ bout, T *__restrict__ fa, const StreamBox &sb,
                                                const T *__restrict__ dtp, int wave)
{
    const int jt = wave % sb.ntj, r = wave / sb.ntj, k = r % sb.nk, seg = r / sb.nk;
    const int j = jt * 64 + (int)lane_id();
    const bool valid = j < sb.nj;
    const int jj = valid ? j : sb.nj - 1;
    const int jm = jj == 0 ? sb.nj - 1 : jj - 1;
    const long N = sb.N;
    const T *rho = u, *mom = u + Comp<FR>::u[1] * N;
    const T *b2 = bin + Comp<FR>::b[1] * N;
    T *b2o = bout + Comp<FR>::b[1] * N;
    const T dt = *dtp;
    const int s0 = seg * sb.S, s1 = min(s0 + sb.S, sb.ni);
    const long kb = (long)k * sb.sk;
    TvdbStream<T> st;
    // the operands of row t+1 are loaded while row t is computed
    long ro = (long)wrap_row(s0 - 4, sb.ni) * sb.si + kb;
    T ma = mom[ro + jm], mb = mom[ro + jj], ra = rho[ro + jm], rb = rho[ro + jj], bt = b2[ro + jj];
    for (int t = s0 - 4; t < s1 + 4; t++) {
        const T cma = ma, cmb = mb, cra = ra, crb = rb, cbt = bt;
        if (t + 1 < s1 + 4) {
            ro = (long)wrap_row(t + 1, sb.ni) * sb.si + kb;
            ma = mom[ro + jm];
            mb = mom[ro + jj];
            ra = rho[ro + jm];
            rb = rho[ro + jj];
            bt = b2[ro + jj];
        }
        T bnew, F;
        if (t < s0 + 3)
            st.push_fill(cma, cmb, cra, crb, cbt, dt);
        else
            st.push(cma, cmb, cra, crb, cbt, dt, bnew, F);
        if (t >= s0 + 4 && valid) {
            const long oo = (long)(t - 4) * sb.si + kb + jj;
            b2o[oo] = bnew;
            fa[oo] = F;
        }
    }
}

// pass B: b3 along i with the k-face velocity; b1 from FA_j, FA_{j+1} (fa)
// and FB_k, FB_{k+1}.  A lane holds the pencils of two neighbouring planes
// k0 and k0+1 (one where k0 is the last plane of an odd count): three
// streams at the faces k0, k0+1, k0+2 give both pencils' fluxes, the middle
// one shared.
template <typename T, int FR>
__device__ __forceinline__ void advect_stream_b(const T *__restrict__ u, const T *__restrict__ bin,
                                                T *__restrict__ bout, const T *__restrict__ fa, const StreamBox &sb,
                                                const T *__restrict__ dtp, int wave)
{
    const int nk2 = (sb.nk + 1) / 2;
    const int jt = wave % sb.ntj, r = wave / sb.ntj, kq = r % nk2, seg = r / nk2;
    const int j = jt * 64 + (int)lane_id();
    const bool valid = j < sb.nj;
    const int jj = valid ? j : sb.nj - 1;
    const int jp = jj + 1 == sb.nj ? 0 : jj + 1;
    const long N = sb.N;
    const T *rho = u, *mom = u + Comp<FR>::u[1] * N;
    const T *b1 = bin + Comp<FR>::b[0] * N, *b3 = bin + Comp<FR>::b[2] * N;
    T *b1o = bout + Comp<FR>::b[0] * N, *b3o = bout + Comp<FR>::b[2] * N;
    const T dt = *dtp;
    const int s0 = seg * sb.S, s1 = min(s0 + sb.S, sb.ni);
    const int k0 = 2 * kq;
    const bool two = k0 + 1 < sb.nk;  // uniform over the wave
    const int km = k0 == 0 ? sb.nk - 1 : k0 - 1;
    const int k1 = k0 + 1 == sb.nk ? 0 : k0 + 1;
    const int k2 = k1 + 1 == sb.nk ? 0 : k1 + 1;
    const long kmb = (long)km * sb.sk, k0b = (long)k0 * sb.sk, k1b = (long)k1 * sb.sk, k2b = (long)k2 * sb.sk;
    TvdbStream<T> sa, sbk, sc;  // faces k0, k0+1, k0+2
    // the operands of row t+1 are loaded while row t is computed
    struct Row {
        T mm, rm, m0, r0, m1, r1, m2, r2, b30, b31, b32;
    };
    auto load_row = [&](int t, Row &q) {
        const long ri = (long)wrap_row(t, sb.ni) * sb.si + jj;
        q.mm = mom[ri + kmb];
        q.rm = rho[ri + kmb];
        q.m0 = mom[ri + k0b];
        q.r0 = rho[ri + k0b];
        q.m1 = mom[ri + k1b];
        q.r1 = rho[ri + k1b];
        q.b30 = b3[ri + k0b];
        q.b31 = b3[ri + k1b];
        if (two) {
            q.m2 = mom[ri + k2b];
            q.r2 = rho[ri + k2b];
            q.b32 = b3[ri + k2b];
        } else {
            q.m2 = q.r2 = q.b32 = (T)0;
        }
    };
    Row nx;
    load_row(s0 - 4, nx);
    for (int t = s0 - 4; t < s1 + 4; t++) {
        const Row q = nx;
        if (t + 1 < s1 + 4) load_row(t + 1, nx);
        T b3n0, F0, b3n1, F1;
        T F2 = 0;
        if (t < s0 + 3) {
            sa.push_fill(q.mm, q.m0, q.rm, q.r0, q.b30, dt);
            sbk.push_fill(q.m0, q.m1, q.r0, q.r1, q.b31, dt);
            if (two) sc.push_fill(q.m1, q.m2, q.r1, q.r2, q.b32, dt);
        } else {
            sa.push(q.mm, q.m0, q.rm, q.r0, q.b30, dt, b3n0, F0);
            sbk.push(q.m0, q.m1, q.r0, q.r1, q.b31, dt, b3n1, F1);
            if (two) {
                T b3x;
                sc.push(q.m1, q.m2, q.r1, q.r2, q.b32, dt, b3x, F2);
            }
        }
        if (t >= s0 + 4 && valid) {
            const long row = (long)(t - 4) * sb.si;
            const long o0 = row + k0b;
            T v = b1[o0 + jj];
            v = v - fa[o0 + jj];
            v = v + fa[o0 + jp];
            v = v - F0;
            v = v + F1;
            b1o[o0 + jj] = v;
            b3o[o0 + jj] = b3n0;
            if (two) {
                const long o1 = row + k1b;
                T w = b1[o1 + jj];
                w = w - fa[o1 + jj];
                w = w + fa[o1 + jp];
                w = w - F1;
                w = w + F2;
                b1o[o1 + jj] = w;
                b3o[o1 + jj] = b3n1;
            }
        }
    }
}

// Both passes in one stream (used when the pencils of a row fit one
// workgroup, nj <= 1024).  A workgroup holds the ntj waves of one plane pair
// and segment, wave w covering pencils 64w .. 64w+63; a lane streams the
// pencil j of planes k0 and k0+1 through five tvdb streams: pass A at k0 and
// k0+1, pass B at the faces k0, k0+1, k0+2.  The b1 update takes FA of pencil
// j+1 from lane l+1, or, across a wave boundary and for the periodic wrap,
// from lane 0 of the neighbouring wave through LDS (xch: 2 buffers x 16
// waves x 2 planes), one barrier per row.  FA never reaches global memory and
// every operand is read once per plane pair.
template <typename T, int FR>
__device__ __forceinline__ void advect_stream_ab(const T *__restrict__ u, const T *__restrict__ bin,
                                                 T *__restrict__ bout, const StreamBox &sb,
                                                 const T *__restrict__ dtp, int kq, int seg, int w, T *xch)
{
    const int j = w * 64 + (int)lane_id();
    const bool valid = j < sb.nj;
    const int jj = valid ? j : sb.nj - 1;
    const int jm = jj == 0 ? sb.nj - 1 : jj - 1;
    const int jp = jj + 1 == sb.nj ? 0 : jj + 1;
    // FA of pencil jp: lane jp & 63 of wave jp >> 6 (lane 0 when another wave)
    const bool remote = (jp >> 6) != w;
    const int src = jp & 63;
    const int wr = jp >> 6;
    const long N = sb.N;
    const T *rho = u, *mom = u + Comp<FR>::u[1] * N;
    const T *b1 = bin + Comp<FR>::b[0] * N, *b2 = bin + Comp<FR>::b[1] * N, *b3 = bin + Comp<FR>::b[2] * N;
    T *b1o = bout + Comp<FR>::b[0] * N, *b2o = bout + Comp<FR>::b[1] * N, *b3o = bout + Comp<FR>::b[2] * N;
    const T dt = *dtp;
    const int s0 = seg * sb.S, s1 = min(s0 + sb.S, sb.ni);
    const int k0 = 2 * kq;
    const bool two = k0 + 1 < sb.nk;  // uniform over the workgroup
    const int km = k0 == 0 ? sb.nk - 1 : k0 - 1;
    const int k1 = k0 + 1 == sb.nk ? 0 : k0 + 1;
    const int k2 = k1 + 1 == sb.nk ? 0 : k1 + 1;
    const long kmb = (long)km * sb.sk, k0b = (long)k0 * sb.sk, k1b = (long)k1 * sb.sk, k2b = (long)k2 * sb.sk;
    const bool multi = sb.ntj > 1;
    TvdbStream<T> a0, a1, fb0, fb1, fb2;  // pass A at k0, k1; pass B at the faces k0, k1, k2
    struct Row {
        T mm0, m0, rm0, r0, b20;  // pencils j-1, j of plane k0; b2 at k0
        T mm1, m1, rm1, r1, b21;  // the same at k1
        T mk, rk, m2, r2;         // pencil j of planes k0-1 and k0+2
        T b30, b31, b32;          // b3 at k0, k1, k2
    };
    auto load_row = [&](int t, Row &q) {
        const long ri = (long)wrap_row(t, sb.ni) * sb.si;
        const long a = ri + k0b, c = ri + k1b;
        q.mm0 = mom[a + jm];
        q.m0 = mom[a + jj];
        q.rm0 = rho[a + jm];
        q.r0 = rho[a + jj];
        q.b20 = b2[a + jj];
        q.mk = mom[ri + kmb + jj];
        q.rk = rho[ri + kmb + jj];
        q.b30 = b3[a + jj];
        q.b31 = b3[c + jj];
        q.m1 = mom[c + jj];
        q.r1 = rho[c + jj];
        q.mm1 = mom[c + jm];
        q.rm1 = rho[c + jm];
        q.b21 = b2[c + jj];
        q.m2 = mom[ri + k2b + jj];
        q.r2 = rho[ri + k2b + jj];
        q.b32 = b3[ri + k2b + jj];
    };
    Row nx;
    load_row(s0 - 4, nx);
    // the five pushes of row t; both pass-A streams and all three face
    // streams run also for the last plane of an odd count (two false),
    // whose second-plane results are dropped, so that the row is straight
    // line code
    T fa0, fa1, f0, f1, f2, b2n0, b2n1, b3n0, b3n1;
    auto push_row = [&](const Row &q) {
        T b3x;
        a0.push(q.mm0, q.m0, q.rm0, q.r0, q.b20, dt, b2n0, fa0);
        a1.push(q.mm1, q.m1, q.rm1, q.r1, q.b21, dt, b2n1, fa1);
        fb0.push(q.mk, q.m0, q.rk, q.r0, q.b30, dt, b3n0, f0);
        fb1.push(q.m0, q.m1, q.r0, q.r1, q.b31, dt, b3n1, f1);
        fb2.push(q.m1, q.m2, q.r1, q.r2, q.b32, dt, b3x, f2);
    };
    // the first eight rows fill the windows (the limiter from the eighth on)
    auto fill_row = [&](const Row &q) {
        a0.push_fill(q.mm0, q.m0, q.rm0, q.r0, q.b20, dt);
        a1.push_fill(q.mm1, q.m1, q.rm1, q.r1, q.b21, dt);
        fb0.push_fill(q.mk, q.m0, q.rk, q.r0, q.b30, dt);
        fb1.push_fill(q.m0, q.m1, q.r0, q.r1, q.b31, dt);
        fb2.push_fill(q.m1, q.m2, q.r1, q.r2, q.b32, dt);
    };
    for (int t = s0 - 4; t < s0 + 3; t++) {
        const Row q = nx;
        load_row(t + 1, nx);
        fill_row(q);
    }
    {
        const Row q = nx;
        load_row(s0 + 4, nx);
        push_row(q);
    }
    // b1 of the output row, loaded one row ahead (rows wrap, so the load
    // after the last row stays inside the plane)
    auto load_b1 = [&](int r, T &c0, T &c1) {
        const long o = (long)wrap_row(r, sb.ni) * sb.si + jj;
        c0 = b1[o + k0b];
        c1 = b1[o + k1b];
    };
    T c10, c11;
    load_b1(s0, c10, c11);
    // row t >= s0 + 4 emits row t - 4; the operands of row t + 1 and b1 of
    // row t - 3 are loaded while row t is computed
    auto row = [&](int t) {
        const Row q = nx;
        const T d10 = c10, d11 = c11;
        load_row(t + 1, nx);
        load_b1(t - 3, c10, c11);
        push_row(q);
        // FA of pencil j+1 at the output row
        T n0 = shfl_from(fa0, src), n1 = shfl_from(fa1, src);
        if (multi) {
            T *x = xch + (t & 1) * 32;
            if (lane_id() == 0) {
                x[w * 2] = fa0;
                x[w * 2 + 1] = fa1;
            }
            __syncthreads();
            if (remote) {
                n0 = x[wr * 2];
                n1 = x[wr * 2 + 1];
            }
        }
        if (valid) {
            const long o = (long)(t - 4) * sb.si + jj;
            T v = d10;
            v = v - fa0;
            v = v + n0;
            v = v - f0;
            v = v + f1;
            b1o[o + k0b] = v;
            b2o[o + k0b] = b2n0;
            b3o[o + k0b] = b3n0;
            if (two) {
                T y = d11;
                y = y - fa1;
                y = y + n1;
                y = y - f1;
                y = y + f2;
                b1o[o + k1b] = y;
                b2o[o + k1b] = b2n1;
                b3o[o + k1b] = b3n1;
            }
        }
    };
    for (int t = s0 + 4; t < s1 + 4; t++) row(t);
}

}  // namespace mhd
