// Batched block-AMR axis pass (one launch covers every leaf of every level).
//
// A work item is (leaf, group of P pencils along the pass axis D); the whole
// pencil of B cells plus the 5 ghost layers on each side is staged in LDS
// (interior cells from the pool; ghosts from a same-level neighbour's cells
// where there is one, else from the leaf's face slabs), the face
// fluxes come from tile_faces (the uniform pass's arithmetic), and the
// residual is accumulated as in block_rhs (amr.rs:1224-1299):
//   rhs -= (F_{i+1/2} - F_{i-1/2}) / dx_level
// in the reference's axis order; the two boundary face fluxes of every pencil
// are stored for refluxing.  The grid is persistent: workgroups stride over
// the work items, so one launch serves any number of leaves.
#pragma once
#include "em_amr_types.h"
#include "em_faces.h"

namespace em {

namespace EM_VARIANT {

// LDS of one work item (whole pencils of a B-cell block, P of them)
template <class Real, int MODE, int B, int P>
struct AmrLds {
    static constexpr bool HYB = MODE == MODE_HYBRID;
    Real sw[5][B + 10][P];
    Real sre[B + 10][P];  // rho_e of each cell
    Real sm[5][B + 5][P];
    Real ss[2][HYB ? B + 10 : 1][P];
    Real sth[HYB ? B + 1 : 1][P];
};

// Work item w of one batched AMR pass: the pencil group w % ngroup of leaf
// leaf0 + w / ngroup (every thread of the workgroup calls it; it returns
// after a barrier).  Counters of shock-path faces, fallbacks and limited
// faces are added to the calling thread's n_*.
template <int D, int MODE, int B, int P, int NT, class Real>
__device__ __forceinline__ void amr_item(const AmrPassArgsT<Real>& pa, long long w, AmrLds<Real, MODE, B, P>& L,
                                         unsigned& n_shock, unsigned& n_fb, unsigned& n_lim) {
    constexpr int C = B + 10;
    constexpr bool HYB = MODE == MODE_HYBRID;
    constexpr int E0 = D == 0 ? 1 : 0;
    constexpr int E1 = D == 2 ? 1 : 2;
    auto& sw = L.sw;
    auto& sre = L.sre;
    auto& sm = L.sm;
    auto& ss = L.ss;
    auto& sth = L.sth;
    const AmrDevT<Real>& a = pa.a;
    const int tid = threadIdx.x;
    const int be1 = a.bd[E1];
    const int nt = a.bcells / B;  // pencils of a block along D
    const int ngh = a.ng;
    const EosT<Real> e = a.eos;
    const double dtv = pa.stage_dt_host ? pa.stage_dt : a.sc->dt;
    const int li = pa.leaf0 + (int)(w / pa.ngroup);
    const int gp = (int)(w % pa.ngroup);
    const long long slot = pa.leaves[li];
    const int lvl = pa.leaf_level[li];
    const Real dx = a.dx[lvl][D];
    const int t0 = gp * P;
    const int Pv = min(P, nt - t0);
    // ---- load the pencils (interior and same-level ghosts from the pool, others from the slabs)
    for (int it = tid; it < C * P; it += NT) {
        int c, p;
        if (pa.dfast) {
            c = it % C;
            p = it / C;
        } else {
            p = it % P;
            c = it / P;
        }
        if (p >= Pv) continue;
        const int t = t0 + p;
        const int x = c - 5;
        Real u[5], wv[5];
        if (x >= 0 && x < B) {
            int cc[3];
            cc[D] = x;
            cc[E0] = t / be1;
            cc[E1] = t % be1;
            const long long cell = ((long long)cc[0] * a.bd[1] + cc[1]) * a.bd[2] + cc[2];
#pragma unroll
            for (int q = 0; q < 5; ++q) u[q] = a.U[amr_u(a, q, slot, cell)];
        } else {
            const int s = 2 * D + (x < 0 ? 0 : 1);
            const int src = pa.gsrc ? pa.gsrc[slot * 6 + s] : -1;
            if (src >= 0) {
                // a same-level neighbour's cells, as its slab copy holds them
                int cc[3];
                cc[D] = x < 0 ? x + B : x - B;
                cc[E0] = t / be1;
                cc[E1] = t % be1;
                const long long cell = ((long long)cc[0] * a.bd[1] + cc[1]) * a.bd[2] + cc[2];
#pragma unroll
                for (int q = 0; q < 5; ++q) u[q] = a.U[amr_u(a, q, src, cell)];
            } else {
                const int layer = x < 0 ? x + ngh : x - B;
#pragma unroll
                for (int q = 0; q < 5; ++q) u[q] = a.G[amr_g(a, q, slot, s, layer, t)];
            }
        }
        to_prim(e, u, wv);
#pragma unroll
        for (int q = 0; q < 5; ++q) sw[q][c][p] = wv[q];
        sre[c][p] = eos_rho_e(e, wv[0], wv[4]);
        if constexpr (HYB) {
            ss[0][c][p] = log(fmax(wv[0], Real(1e-300)));
            ss[1][c][p] = log(fmax(wv[4] + e.p_inf, Real(1e-300)));
        }
    }
    __syncthreads();
    // amr.rs:1254; formed in double from the device dt and rounded to Real
    const double lam = pa.ndim_pos > 0.0 ? pa.ndim_pos * dtv / (double)dx : 0.0;
    tile_faces<D, MODE, B, P, NT>(sw, sre, sm, ss, sth, B, Pv, 0, B, e, Real(lam), pa.tol, n_shock, n_fb, n_lim);
    // ---- residual and boundary face fluxes
    for (int it = tid; it < B * P; it += NT) {
        int c, p;
        if (pa.dfast) {
            c = it % B;
            p = it / B;
        } else {
            p = it % P;
            c = it / P;
        }
        if (p >= Pv) continue;
        const int t = t0 + p;
        int cc[3];
        cc[D] = c;
        cc[E0] = t / be1;
        cc[E1] = t % be1;
        const long long cell = ((long long)cc[0] * a.bd[1] + cc[1]) * a.bd[2] + cc[2];
#pragma unroll
        for (int q = 0; q < 5; ++q) {
            const long long ix = amr_u(a, q, slot, cell);
            const Real prev = pa.first ? Real(0.0) : a.R[ix];
            a.R[ix] = prev - (sm[q][c + 1][p] - sm[q][c][p]) / dx;
        }
    }
    for (int it = tid; it < 2 * P; it += NT) {
        const int p = it % P, side = it / P;
        if (p >= Pv) continue;
        const int t = t0 + p;
#pragma unroll
        for (int q = 0; q < 5; ++q) a.F[amr_f(a, q, slot, 2 * D + side, t)] = sm[q][side ? B : 0][p];
    }
    __syncthreads();
}

template <int D, int MODE, int B, int P, int NT, class Real>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(EM_WAVES_PER_EU))) void amr_pass(AmrPassArgsT<Real> pa) {
    __shared__ AmrLds<Real, MODE, B, P> L;
    __shared__ unsigned int scnt[3];

    const AmrDevT<Real>& a = pa.a;
    if (a.sc->active == 0) return;
    const int tid = threadIdx.x;
    if (tid < 3) scnt[tid] = 0;
    const long long nwork = (long long)pa.nleaves * pa.ngroup;
    unsigned int n_shock = 0, n_fb = 0, n_lim = 0;
    // plain stride: the chunked XCD order (xcd_stride) measured no faster here
    for (long long w = blockIdx.x; w < nwork; w += gridDim.x)
        amr_item<D, MODE, B, P, NT>(pa, w, L, n_shock, n_fb, n_lim);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        n_shock += __shfl_xor(n_shock, off);
        n_fb += __shfl_xor(n_fb, off);
        n_lim += __shfl_xor(n_lim, off);
    }
    if ((tid & 63) == 0) {
        if (n_shock) atomicAdd(&scnt[0], n_shock);
        if (n_fb) atomicAdd(&scnt[1], n_fb);
        if (n_lim) atomicAdd(&scnt[2], n_lim);
    }
    __syncthreads();
    if (tid == 0) {
        const int slotc = blockIdx.x % kSlots;
        if (scnt[0]) atomicAdd(&a.sc->counters[0][slotc][0], (unsigned long long)scnt[0]);
        if (scnt[1]) atomicAdd(&a.sc->counters[1][slotc][0], (unsigned long long)scnt[1]);
        if (scnt[2]) atomicAdd(&a.sc->counters[2][slotc][0], (unsigned long long)scnt[2]);
    }
}

// Adds a workgroup's counters to hierarchy `sc` (every thread calls it; the
// n_* are reset).
__device__ __forceinline__ void amr_flush_counters(DevScalars* sc, unsigned* scnt, unsigned& n_shock, unsigned& n_fb,
                                                   unsigned& n_lim) {
    const int tid = threadIdx.x;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        n_shock += __shfl_xor(n_shock, off);
        n_fb += __shfl_xor(n_fb, off);
        n_lim += __shfl_xor(n_lim, off);
    }
    if (tid < 3) scnt[tid] = 0;
    __syncthreads();
    if ((tid & 63) == 0) {
        if (n_shock) atomicAdd(&scnt[0], n_shock);
        if (n_fb) atomicAdd(&scnt[1], n_fb);
        if (n_lim) atomicAdd(&scnt[2], n_lim);
    }
    __syncthreads();
    if (tid == 0) {
        const int slotc = blockIdx.x % kSlots;
        if (scnt[0]) atomicAdd(&sc->counters[0][slotc][0], (unsigned long long)scnt[0]);
        if (scnt[1]) atomicAdd(&sc->counters[1][slotc][0], (unsigned long long)scnt[1]);
        if (scnt[2]) atomicAdd(&sc->counters[2][slotc][0], (unsigned long long)scnt[2]);
    }
    __syncthreads();
    n_shock = n_fb = n_lim = 0;
}

// The batched pass over the leaves of several hierarchies, whose blocks may
// differ in size (8, 16 or 32 cells a side), in one persistent launch: work
// item w belongs to the hierarchy h with prefix[h] <= w < prefix[h + 1],
// and the item body is the single-hierarchy pass's (amr_item) for that
// hierarchy's block size, in the LDS of a union sized for the largest.
// Each hierarchy brings its own pool, tables, dt, counters and options.
// Three waves per SIMD (168 VGPRs): the switch over three block sizes needs
// more registers than the single-size pass, which spills 200+ bytes per lane
// at four waves and 48-68 at three.
template <int D, int MODE>
__global__ __launch_bounds__(kAmrMultiThreads) __attribute__((amdgpu_waves_per_eu(3))) void amr_multi(AmrMultiArgs m) {
    constexpr int NT = kAmrMultiThreads;
    __shared__ union {
        AmrLds<double, MODE, 8, 16> b8;
        AmrLds<double, MODE, 16, 12> b16;
        AmrLds<double, MODE, 32, 6> b32;
    } L;
    __shared__ unsigned int scnt[3];
    unsigned int n_shock = 0, n_fb = 0, n_lim = 0;
    int cur = -1;
    for (long long w = blockIdx.x; w < m.total; w += gridDim.x) {
        int lo = 0, hi = m.n;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (m.prefix[mid] <= w)
                lo = mid;
            else
                hi = mid;
        }
        const AmrPassArgs& pa = m.args[lo];  // read in place (global memory, cached)
        if (lo != cur) {
            if (cur >= 0) amr_flush_counters(m.args[cur].a.sc, scnt, n_shock, n_fb, n_lim);
            cur = lo;
        }
        if (pa.a.sc->active == 0) continue;
        const long long wl = w - m.prefix[lo];
        switch (pa.a.bd[D]) {
            case 8: amr_item<D, MODE, 8, 16, NT>(pa, wl, L.b8, n_shock, n_fb, n_lim); break;
            case 16: amr_item<D, MODE, 16, 12, NT>(pa, wl, L.b16, n_shock, n_fb, n_lim); break;
            default: amr_item<D, MODE, 32, 6, NT>(pa, wl, L.b32, n_shock, n_fb, n_lim); break;
        }
    }
    if (cur >= 0) amr_flush_counters(m.args[cur].a.sc, scnt, n_shock, n_fb, n_lim);
}

template <int D, int MODE>
static hipError_t launch_amr_multi_d(const AmrMultiArgs& m, unsigned grid, hipStream_t s) {
    hipLaunchKernelGGL((amr_multi<D, MODE>), dim3(grid), dim3(kAmrMultiThreads), 0, s, m);
    return hipGetLastError();
}

// Block sizes with a compiled pass: B = 8, 16, 32, in the shapes of
// amr_shape (P pencils, NT threads).
template <int D, int MODE, class Real>
static hipError_t launch_amr_b(int b, int shape, const AmrPassArgsT<Real>& pa, unsigned grid, hipStream_t s) {
#define EM_AMR_CASE(B_, P_, NT_) \
    hipLaunchKernelGGL((amr_pass<D, MODE, B_, P_, NT_, Real>), dim3(grid), dim3(NT_), 0, s, pa)
    switch (b * 16 + shape) {
        case 8 * 16 + 0: EM_AMR_CASE(8, 8, 128); break;
        case 8 * 16 + 1: EM_AMR_CASE(8, 16, 256); break;
        case 8 * 16 + 2: EM_AMR_CASE(8, 32, 512); break;
        case 8 * 16 + 3: EM_AMR_CASE(8, 64, 512); break;
        case 16 * 16 + 0: EM_AMR_CASE(16, 16, 384); break;
        case 16 * 16 + 1: EM_AMR_CASE(16, 32, 512); break;
        case 32 * 16 + 0: EM_AMR_CASE(32, 8, 320); break;
        case 32 * 16 + 1: EM_AMR_CASE(32, 16, 512); break;
        default: return hipErrorInvalidValue;
    }
#undef EM_AMR_CASE
    return hipGetLastError();
}

}  // namespace EM_VARIANT

}  // namespace em
