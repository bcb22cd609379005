// Face fluxes of a tile of P pencils held in LDS: the hybrid blend weight
// (scheme.rs:171-183), the WCNS midpoints (scheme.rs:286-333, each computed
// once), the S6 midpoint-to-node combination with the KEEP-PE central flux
// and theta blend (scheme.rs:188-276), and the positivity limiter
// (scheme.rs:395-464).  Shared by the uniform-grid axis pass and the batched
// AMR pass so both evaluate identical arithmetic.
//
// On entry (after a barrier): sw[k][c][p] holds the primitive state of the
// tile cells c = 0 .. Tv+9 (cell c0-5+c of the pencil), sre[c][p] its
// rho_e = eos_rho_e(rho, p) (used by the central pair fluxes and the
// limiter's conserved states), ss the hybrid sensor.  On exit (after a
// barrier): sm[k][f][p] holds the flux of face f = 0 .. Tv (between cells
// c0+f-1 and c0+f).  Counters of midpoints and faces
// owned by the tile (tiles of one pencil share their halo) are accumulated.
#pragma once
#include "em_device.h"
#include "em_types.h"

namespace em {
namespace EM_VARIANT {

template <int D, int MODE, int T, int P, int NT, class R>
__device__ __forceinline__ void tile_faces(R (&sw)[5][T + 10][P], R (&sre)[T + 10][P],
                                           R (&sm)[5][T + 5][P],
                                           R (&ss)[2][MODE == MODE_HYBRID ? T + 10 : 1][P],
                                           R (&sth)[MODE == MODE_HYBRID ? T + 1 : 1][P], int Tv, int Pv,
                                           int c0, int nD, const EosT<R>& e, R lam, R tol,
                                           unsigned& n_shock, unsigned& n_fb, unsigned& n_lim) {
    constexpr int M = T + 5;
    constexpr int NF = T + 1;
    constexpr bool HYB = MODE == MODE_HYBRID;
    constexpr bool CEN = MODE == MODE_CENTRAL;
    constexpr int FITER = (NF * P + NT - 1) / NT;
    const int tid = threadIdx.x;
    // ---- 2. hybrid blend weights (scheme.rs:171-183)
    if constexpr (HYB) {
        for (int it = tid; it < NF * P; it += NT) {
            const int p = it % P, f = it / P;
            if (p >= Pv || f > Tv) continue;
            R r = R(0.0);
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                R q[6];
#pragma unroll
                for (int j = 0; j < 6; ++j) q[j] = ss[k][f + 2 + j][p];
                r = fmax(fmax(r, d4(q[0], q[1], q[2], q[3], q[4])), d4(q[1], q[2], q[3], q[4], q[5]));
            }
            R x = (r / tol - R(0.5)) / R(0.5);
            if (x < R(0.0)) x = R(0.0);
            if (x > R(1.0)) x = R(1.0);
            sth[f][p] = isfinite(r) ? x * x * (R(3.0) - R(2.0) * x) : R(1.0);
        }
        __syncthreads();
    }

    
    // ---- 3. WCNS midpoints (midpoint m sits between cells c0-3+m and c0-2+m)
    if constexpr (!CEN) {
        for (int it = tid; it < M * P; it += NT) {
            const int p = it % P, m = it / P;
            if (p >= Pv || m >= Tv + 5) continue;
            if constexpr (HYB) {
                bool need = false;
                const int flo = max(0, m - 4), fhi = min(Tv, m);
                for (int f = flo; f <= fhi; ++f) need = need || sth[f][p] > R(0.0);
                if (!need) continue;
            }
            R s[6][5];
#pragma unroll
            for (int j = 0; j < 6; ++j)
#pragma unroll
                for (int k = 0; k < 5; ++k) s[j][k] = sw[k][m + j][p];
            R f[5];
            const bool fb = midpoint_flux<D>(e, s, f);
#pragma unroll
            for (int k = 0; k < 5; ++k) sm[k][m][p] = f[k];
            if (fb) {
                const int j = c0 - 3 + m;  // global midpoint index, owned once per pencil
                const bool own = (j >= c0 && j < c0 + Tv) || (j < 0 && c0 == 0) || (j >= nD && c0 + Tv == nD);
                n_fb += own ? 1u : 0u;
            }
        }
        __syncthreads();
    }

    // ---- 4. face fluxes (face f between cells c0+f-1 and c0+f)
    R FR[FITER][5];
#pragma unroll
    for (int fi = 0; fi < FITER; ++fi) {
        const int it = tid + fi * NT;
        const int p = it % P, f = it / P;
        if (it >= NF * P || p >= Pv || f > Tv) continue;
        // a face shared by two tiles is counted by the tile on its right
        const unsigned own_face = (f < Tv || c0 + Tv == nD) ? 1u : 0u;
        R theta;
        if constexpr (MODE == MODE_WENO)
            theta = R(1.0);
        else if constexpr (CEN)
            theta = R(0.0);
        else
            theta = sth[f][p];
        R fo[5] = {R(0.0), R(0.0), R(0.0), R(0.0), R(0.0)};
        R cen[5] = {R(0.0), R(0.0), R(0.0), R(0.0), R(0.0)};
        if (theta < R(1.0)) {
            // KEEP-PE 6th-order central flux, scheme.rs:188-214
            const R a6[3] = {R(A6_0), R(A6_1), R(A6_2)};
#pragma unroll
            for (int li = 0; li < 3; ++li) {
                const int lag = li + 1;
#pragma unroll
                for (int mm = 0; mm < lag; ++mm) {
                    R wa[5], wb[5], ft[5];
#pragma unroll
                    for (int k = 0; k < 5; ++k) {
                        wa[k] = sw[k][f + 4 - mm][p];
                        wb[k] = sw[k][f + 4 - mm + lag][p];
                    }
                    split_flux_re<D>(wa, wb, sre[f + 4 - mm][p], sre[f + 4 - mm + lag][p], ft);
#pragma unroll
                    for (int k = 0; k < 5; ++k) cen[k] += R(2.0) * a6[li] * ft[k];
                }
            }
        }
        if (theta == R(0.0)) {
#pragma unroll
            for (int k = 0; k < 5; ++k) fo[k] = cen[k];
        } else {
            n_shock += own_face;
            // F = sum_s S6[s] (F_{i+s} + F_{i-s}), order s=0, i+1, i-1, i+2, i-2
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                R v = R(0.0);
                v += R(S6_0) * sm[k][f + 2][p];
                v += R(S6_1) * sm[k][f + 3][p];
                v += R(S6_1) * sm[k][f + 1][p];
                v += R(S6_2) * sm[k][f + 4][p];
                v += R(S6_2) * sm[k][f][p];
                fo[k] = v;
            }
            if (theta < R(1.0)) {
#pragma unroll
                for (int k = 0; k < 5; ++k) fo[k] = theta * fo[k] + (R(1.0) - theta) * cen[k];
            }
        }
        if (lam > R(0.0)) {
            // positivity limiter, scheme.rs:401-444
            R wl[5], wr[5], ul[5], ur[5], fl[5], fr[5];
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                wl[k] = sw[k][f + 4][p];
                wr[k] = sw[k][f + 5][p];
            }
            to_cons_re(wl, sre[f + 4][p], ul);
            to_cons_re(wr, sre[f + 5][p], ur);
            phys_flux<D>(wl, ul, fl);
            phys_flux<D>(wr, ur, fr);
            const R tl = R(2.0) * lam;
            auto ok = [&](const R* fx) {
                R qa[5], qb[5], pa[5], pb[5];
#pragma unroll
                for (int k = 0; k < 5; ++k) {
                    qa[k] = ul[k] - tl * (fx[k] - fl[k]);
                    qb[k] = ur[k] + tl * (fx[k] - fr[k]);
                }
                to_prim(e, qa, pa);
                to_prim(e, qb, pb);
                return admissible(e, pa[0], pa[4]) && admissible(e, pb[0], pb[4]);
            };
            bool fin = true;
#pragma unroll
            for (int k = 0; k < 5; ++k) fin = fin && isfinite(fo[k]);
            if (!(fin && ok(fo))) {
                constexpr int un = 1 + D;
                const R al = fmax(fabs(wl[un]) + sound_speed(e, wl[0], wl[4]),
                                       fabs(wr[un]) + sound_speed(e, wr[0], wr[4]));
                R flf[5], fh[5];
#pragma unroll
                for (int k = 0; k < 5; ++k) {
                    flf[k] = R(0.5) * (fl[k] + fr[k]) - R(0.5) * al * (ur[k] - ul[k]);
                    fh[k] = fo[k];
                }
                n_lim += own_face;
                R lo = R(0.0), hi = R(1.0);
                for (int itb = 0; itb < 30; ++itb) {
                    const R mid = R(0.5) * (lo + hi);
                    R bl[5];
#pragma unroll
                    for (int k = 0; k < 5; ++k) bl[k] = isfinite(fh[k]) ? mid * fh[k] + (R(1.0) - mid) * flf[k] : flf[k];
                    if (ok(bl))
                        lo = mid;
                    else
                        hi = mid;
                }
#pragma unroll
                for (int k = 0; k < 5; ++k) fo[k] = isfinite(fh[k]) ? lo * fh[k] + (R(1.0) - lo) * flf[k] : flf[k];
            }
        }
#pragma unroll
        for (int k = 0; k < 5; ++k) FR[fi][k] = fo[k];
    }
    __syncthreads();  // every midpoint read is done: reuse sm for face fluxes
#pragma unroll
    for (int fi = 0; fi < FITER; ++fi) {
        const int it = tid + fi * NT;
        const int p = it % P, f = it / P;
        if (it >= NF * P || p >= Pv || f > Tv) continue;
#pragma unroll
        for (int k = 0; k < 5; ++k) sm[k][f][p] = FR[fi][k];
    }
    __syncthreads();

}

}  // namespace EM_VARIANT
}  // namespace em
