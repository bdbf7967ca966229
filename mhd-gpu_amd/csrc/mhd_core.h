// Per-pencil compute core of the sweeps, shared by every kernel layout:
// a pencil of n cells lives in one wavefront, lane l < L holding cells
// l*K .. l*K+K-1 in registers; lanes >= L shadow lane l % L.  Neighbours along
// the pencil come from lane shuffles, the pencil max from a wave reduction.
// The lane geometry LN is LanesFull (n = L*K), LanesRagged (the last lane
// holds fewer than K cells; its other cells stay out of every reduction) or
// LanesGroup<G> (64/G pencils of G*K cells share the wave, each on its own
// aligned group of G lanes; shifts and the max stay inside the group).
#pragma once
#include <type_traits>

#include "mhd_common.h"

namespace mhd {

// Unnormalised fluxes v and the pencil's freezing speed c (mhdflux.c:21-69),
// then v/c.
template <typename T, int K, class LN>
__device__ __forceinline__ void mhdflux_pencil(const T (&uu)[K][5], const T (&bx)[K][3], T (&v)[K][5], T &c,
                                               const LN &ln)
{
    constexpr bool whole = !std::is_same<LN, LanesRagged>::value;
    T cl = (T)0;
    if constexpr (sizeof(T) == 4) {
        // the divisions by u1 through one reciprocal per cell (exact while no
        // operand scaling would trigger, as for the normalisation below);
        // the wave recomputes its cells with true divisions if any operand
        // of any lane is out of range
        bool uok = true;
        float vmax = 0.0f;
        int emin = 0;
#pragma unroll
        for (int q = 0; q < K; q++) {
            T ci;
            if constexpr (whole) {
                mhdflux_cell_bulk(uu[q], bx[q], v[q], ci, uok, vmax, emin);
                cl = fmax(cl, ci);
            } else {
                bool ok1 = true;
                float vm1 = 0.0f;
                int em1 = 0;
                mhdflux_cell_bulk(uu[q], bx[q], v[q], ci, ok1, vm1, em1);
                if (ln.valid(q)) {
                    uok = uok && ok1;
                    vmax = fmaxf(vmax, vm1);
                    emin = min(emin, em1);
                    cl = fmax(cl, ci);
                }
            }
        }
        if (!__all(uok && vmax < 0x1p60f && emin >= -60)) {
            cl = (T)0;
#pragma unroll
            for (int q = 0; q < K; q++) {
                T ci;
                mhdflux_cell<T>(uu[q], bx[q], v[q], ci);
                if (ln.valid(q)) cl = fmax(cl, ci);
            }
        }
    } else {
#pragma unroll
        for (int q = 0; q < K; q++) {
            T ci;
            mhdflux_cell<T>(uu[q], bx[q], v[q], ci);
            if (ln.valid(q)) cl = fmax(cl, ci);
        }
    }
    c = pencil_max(cl, ln);
    if (c > (T)0) {
        if constexpr (sizeof(T) == 4) {
            // v/c for the 5K fluxes of the lane: c is uniform over the
            // pencil's lanes, so its reciprocal is formed once and each
            // quotient takes a product and one residual correction, which
            // equals the IEEE quotient over the ranges below (Recip).
            // The operand ranges are checked in bulk: c finite with exponent
            // in [-30, 30]; the largest |v| below 2^60 (which also rules out
            // infinities) and the smallest exponent at least -60 (zero has
            // exponent 0 and divides exactly, up to the sign of the zero; NaN
            // gives NaN either way).  The wave falls back to true divisions
            // if any lane is outside.
            const int ec = __builtin_amdgcn_frexp_expf(c);
            const bool cok = ec >= -30 && ec <= 30 && c < __builtin_inff();
            float vmax = 0.0f;
            int emin = 0;
#pragma unroll
            for (int q = 0; q < K; q++)
#pragma unroll
                for (int m = 0; m < 5; m++) {
                    if (ln.valid(q)) {
                        vmax = fmaxf(vmax, fabsf(v[q][m]));
                        emin = min(emin, __builtin_amdgcn_frexp_expf(v[q][m]));
                    }
                }
            if (__all(cok && vmax < 0x1p60f && emin >= -60)) {
                const Recip rc = recip_of(c);
#pragma unroll
                for (int q = 0; q < K; q++)
#pragma unroll
                    for (int m = 0; m < 5; m++) v[q][m] = quot_by(v[q][m], rc);
            } else {
#pragma unroll
                for (int q = 0; q < K; q++)
#pragma unroll
                    for (int m = 0; m < 5; m++) v[q][m] = v[q][m] / c;
            }
        } else {
#pragma unroll
            for (int q = 0; q < K; q++)
#pragma unroll
                for (int m = 0; m < 5; m++) v[q][m] = v[q][m] / c;
        }
    }
}

// The products c*w of tvd1.c:40,43,58,84 are abXYc calls with b = 0, i.e.
// c*w + 0*w (+ 0): the 0*w term is a zero for finite w but NaN for an
// infinite one, so it is kept, as one fma (exact: 0*w is a zero or NaN,
// and adding a zero to c*w rounds nothing).  Measured cost on MI355X: about
// 2% of a 256^3 step; a wave-uniform test that skipped the term while every
// value of the wave was small enough ran no faster (its branches spilled).
template <typename T, int K> __device__ __forceinline__ void cprod(T c, const T (&w)[K], T (&o)[K])
{
#pragma unroll
    for (int q = 0; q < K; q++) o[q] = fma((T)0, w[q], c * w[q]);
}

// tvd1 on a pencil held K cells per lane; uu is updated in place.
template <typename T, int K, class LN>
__device__ __forceinline__ void tvd1_pencil(T (&uu)[K][5], const T (&bx)[K][3], T dt, const LN &ln)
{
    const T h = (T)(-1.0 * (double)dt / 2.0);
    const T mdt = (T)(-1.0 * (double)dt);
    T v[K][5], c;
    T u1[K][5];
    // first half step (tvd1.c:34-49)
    mhdflux_pencil<T, K>(uu, bx, v, c, ln);
#pragma unroll
    for (int m = 0; m < 5; m++) {
        T wr[K], wl[K], fr[K], wln[K], fl[K], fx[K], fxp[K];
#pragma unroll
        for (int q = 0; q < K; q++) {
            wr[q] = uu[q][m] + v[q][m];
            wl[q] = uu[q][m] - v[q][m];
        }
        shift_next<T, K>(wl, wln, ln);
        cprod<T, K>(c, wr, fr);
        cprod<T, K>(c, wln, fl);
#pragma unroll
        for (int q = 0; q < K; q++) fx[q] = (T)0.5 * fr[q] + (T)(-0.5) * fl[q];
        shift_prev<T, K>(fx, fxp, ln);
#pragma unroll
        for (int q = 0; q < K; q++) u1[q][m] = uu[q][m] + h * (fx[q] - fxp[q]);
    }
    // full step with limited fluxes (tvd1.c:52-116)
    mhdflux_pencil<T, K>(u1, bx, v, c, ln);
#pragma unroll
    for (int m = 0; m < 5; m++) {
        // tvd1.c:59-90 forms dfrp = 0.5*fr[i+1] + (-0.5)*fr[i] and
        // dfrm = 0.5*fr[i] + (-0.5)*fr[i-1] (likewise for fl).  In IEEE
        // arithmetic (-0.5)*x == -(0.5*x) and a + (-b) == a - b, so with
        // hfr = 0.5*fr both are differences of neighbouring hfr, and
        // dfrm[i] is dfrp[i-1] bit for bit: one product and one difference
        // per cell instead of four and two.
        T wr[K], wl[K], fr[K], wln[K], fl[K], hr[K], hrn[K], hl[K], hln[K], dfrp[K], dfrm[K], dflp[K], dflm[K],
            fx[K], fxp[K];
#pragma unroll
        for (int q = 0; q < K; q++) {
            wr[q] = u1[q][m] + v[q][m];
            wl[q] = u1[q][m] - v[q][m];
        }
        shift_next<T, K>(wl, wln, ln);
        cprod<T, K>(c, wr, fr);
        cprod<T, K>(c, wln, fl);
#pragma unroll
        for (int q = 0; q < K; q++) {
            hr[q] = (T)0.5 * fr[q];
            hl[q] = (T)0.5 * fl[q];
        }
        shift_next<T, K>(hr, hrn, ln);
        shift_next<T, K>(hl, hln, ln);
#pragma unroll
        for (int q = 0; q < K; q++) {
            dfrp[q] = hrn[q] - hr[q];
            dflp[q] = hl[q] - hln[q];
        }
        shift_prev<T, K>(dfrp, dfrm, ln);
        shift_prev<T, K>(dflp, dflm, ln);
#pragma unroll
        for (int q = 0; q < K; q++) {
            const T dfr = vanleer<T>(dfrp[q], dfrm[q]);
            const T dfl = vanleer<T>(dflp[q], dflm[q]);
            fx[q] = (T)0.5 * (fr[q] - fl[q]) + (T)0.5 * (dfr - dfl);
        }
        shift_prev<T, K>(fx, fxp, ln);
#pragma unroll
        for (int q = 0; q < K; q++) uu[q][m] = uu[q][m] + mdt * (fx[q] - fxp[q]);
    }
}

// tvdb on a pencil held K cells per lane (tvdb.c:10-91): bin -> bnew, and
// F[i] = flux[i-1], the shifted flux the CT update uses.
template <typename T, int K, class LN>
__device__ __forceinline__ void tvdb_pencil(const T (&bin)[K], const T (&vg)[K], T dt, T (&bnew)[K], T (&F)[K],
                                            const LN &ln)
{
    T vgn[K], vh[K], bv[K], bvn[K], f1[K], f1p[K], b1[K], g[K], gn[K], gnn[K], gp[K], fl[K], flp[K];
    shift_next<T, K>(vg, vgn, ln);
#pragma unroll
    for (int q = 0; q < K; q++) {
        vh[q] = (T)0.5 * vg[q] + (T)0.5 * vgn[q];
        bv[q] = bin[q] * vg[q];
    }
    shift_next<T, K>(bv, bvn, ln);
#pragma unroll
    for (int q = 0; q < K; q++) f1[q] = vh[q] > (T)0 ? bv[q] : bvn[q];
    shift_prev<T, K>(f1, f1p, ln);
#pragma unroll
    for (int q = 0; q < K; q++) {
        b1[q] = sub_half<T>(bin[q], (f1[q] - f1p[q]) * dt);
        g[q] = vg[q] * b1[q];
    }
    shift_next<T, K>(g, gn, ln);
    shift_next<T, K>(gn, gnn, ln);
    shift_prev<T, K>(g, gp, ln);
#pragma unroll
    for (int q = 0; q < K; q++) {
        T w, wp, wm;
        if (vh[q] > (T)0) {
            w = g[q];
            wp = half_of<T>(gn[q] - w);
            wm = half_of<T>(w - gp[q]);
        } else {
            w = gn[q];
            wp = half_of<T>(w - gnn[q]);
            wm = half_of<T>(g[q] - w);
        }
        const T dw = vanleer<T>(wm, wp);
        fl[q] = (w + dw) * dt;
    }
    shift_prev<T, K>(fl, flp, ln);
#pragma unroll
    for (int q = 0; q < K; q++) {
        bnew[q] = bin[q] - (fl[q] - flp[q]);
        F[q] = flp[q];
    }
}

// the same for a whole pencil with the neighbour lanes given as indices
template <typename T, int K>
__device__ __forceinline__ void tvd1_pencil(T (&uu)[K][5], const T (&bx)[K][3], T dt, int src_next, int src_prev)
{
    tvd1_pencil<T, K>(uu, bx, dt, LanesFull{src_next, src_prev});
}
template <typename T, int K>
__device__ __forceinline__ void tvdb_pencil(const T (&bin)[K], const T (&vg)[K], T dt, T (&bnew)[K], T (&F)[K],
                                            int src_next, int src_prev)
{
    tvdb_pencil<T, K>(bin, vg, dt, bnew, F, LanesFull{src_next, src_prev});
}

}  // namespace mhd
