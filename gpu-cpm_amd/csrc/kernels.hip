// kernels.hip -- HIP kernels of the MI355X CPM engine (gfx950).
//
// One colour phase of the checkerboard copy-attempt sweep per launch
// (reference: src/kernels2d.cu:44-583, src/kernels3d.cu:51-594), the tally
// rebuild run at push (kernels2d.cu:5-41, kernels3d.cu:3-47) and the
// end-of-run settle of the double-buffered cell tables.
#include <hip/hip_ext.h>

#include "cpm_device.h"

namespace gcpm {

// Philox4x32-10 (Salmon et al., SC'11).  Round r uses key + r * (W0, W1).
__device__ __forceinline__ void philox4x32_10(uint32_t x0, uint32_t x1, uint32_t x2,
                                              uint32_t x3, uint32_t k0, uint32_t k1,
                                              uint32_t w[4]) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * x0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * x2;
        const uint32_t y0 = (uint32_t)(p1 >> 32) ^ x1 ^ k0;
        const uint32_t y2 = (uint32_t)(p0 >> 32) ^ x3 ^ k1;
        x1 = (uint32_t)p1;
        x3 = (uint32_t)p0;
        x0 = y0;
        x2 = y2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    w[0] = x0;
    w[1] = x1;
    w[2] = x2;
    w[3] = x3;
}

// Neighbour k of the 3^D block in (dz, dy, dx) lexicographic order, dx
// fastest; the centre (k = (3^D - 1) / 2) is the site itself.
template <int D>
__device__ __forceinline__ void block_offset(int k, int& ox, int& oy, int& oz) {
    ox = k % 3 - 1;
    oy = (k / 3) % 3 - 1;
    oz = D == 3 ? k / 9 - 1 : 0;
}

template <int D>
__device__ __forceinline__ uint32_t site_index(int x, int y, int z, int lgL) {
    return (uint32_t)x | ((uint32_t)y << lgL) | (D == 3 ? ((uint32_t)z << (2 * lgL)) : 0u);
}

__device__ __forceinline__ void atomic_add_i64(int64_t* p, int64_t v) {
    atomicAdd(reinterpret_cast<unsigned long long*>(p), (unsigned long long)v);
}

// One colour phase.  Thread q owns same-colour sub-block q and makes its u
// copy attempts in order; sub-blocks of one colour are two sub-blocks apart,
// so no attempt reads a site another thread writes in this phase.
template <int D>
__global__ __launch_bounds__(256) void k_sweep_phase(PhaseArgs a) {
    const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const int gsz = gridDim.x * blockDim.x;
    // Fold the previous phase's deltas into W (W is never read in a phase).
    for (int ci = gtid; ci <= a.ncells; ci += gsz) {
        const int64_t dv = a.Dprev[ci];
        if (dv != 0) {
            atomic_add_i64(a.W + ci, dv);
            a.Dprev[ci] = 0;
        }
    }
    if (gtid >= a.nsub) return;

    constexpr int N = D == 3 ? 26 : 8;
    constexpr int NB = D == 3 ? 27 : 9;
    constexpr int CENTRE = NB / 2;
    const int lgL = a.lgL;
    const int Lm = (1 << lgL) - 1;
    const int lc = a.lc;
    const uint32_t cm = (1u << lc) - 1u;
    const int hm = (1 << a.lh) - 1;
    const int bx = 2 * (gtid & hm) + a.cx;
    const int by = 2 * ((gtid >> a.lh) & hm) + a.cy;
    const int bz = D == 3 ? 2 * (gtid >> (2 * a.lh)) + a.cz : 0;
    const int lsb = lgL - lc;  // log2 sub-blocks per axis
    const uint32_t sb = (uint32_t)bx | ((uint32_t)by << lsb) |
                        (D == 3 ? ((uint32_t)bz << (2 * lsb)) : 0u);
    uint32_t* __restrict__ lat = a.lat;
    const DevParams* __restrict__ P = a.prm;

    for (int i = 0; i < a.u; i++) {
        uint32_t w[4];
        philox4x32_10(sb, a.mcs, a.phase, (uint32_t)i, a.k0, a.k1, w);
        const int tx = (bx << lc) + (int)(w[0] & cm);
        const int ty = (by << lc) + (int)((w[0] >> lc) & cm);
        const int tz = D == 3 ? (bz << lc) + (int)((w[0] >> (2 * lc)) & cm) : 0;
        const int dir = (int)(((uint64_t)w[1] * (uint64_t)N) >> 32);
        int ox, oy, oz;
        block_offset<D>(dir < CENTRE ? dir : dir + 1, ox, oy, oz);
        const uint32_t ti = site_index<D>(tx, ty, tz, lgL);
        const uint32_t si = site_index<D>((tx + ox) & Lm, (ty + oy) & Lm, (tz + oz) & Lm, lgL);
        const uint32_t vs = lat[si];
        const uint32_t vt = lat[ti];
        if (vs == vt) continue;  // kernels2d.cu:276

        const uint32_t ts = vs >> 24, tt = vt >> 24;
        int32_t adh = 0;
        int as = 0, at = 0;
#pragma unroll
        for (int k = 0; k < NB; k++) {
            if (k == CENTRE) continue;
            int nx, ny, nz;
            block_offset<D>(k, nx, ny, nz);
            const uint32_t vn = lat[site_index<D>((tx + nx) & Lm, (ty + ny) & Lm, (tz + nz) & Lm, lgL)];
            const uint32_t tn = vn >> 24;
            if (vn != vs) adh += P->J[ts][tn]; else as++;
            if (vn != vt) adh -= P->JS[tt][tn]; else at++;
        }
        int64_t dh = adh;
        const uint32_t cs = vs & kSpinMask, ct = vt & kSpinMask;
        if (cs) {
            const int64_t row = a.R[cs];
            const int64_t A = (int64_t)(uint32_t)row, Pm = row >> 32;
            const int64_t dP = N - 2 * as;
            dh += (int64_t)P->la[ts] * (1 + 2 * A - 2 * (int64_t)P->a0[ts]);
            dh += (int64_t)P->lp[ts] * dP * (dP + 2 * (Pm - P->p0[ts]));
        }
        if (ct) {
            const int64_t row = a.R[ct];
            const int64_t A = (int64_t)(uint32_t)row, Pm = row >> 32;
            const int64_t dP = 2 * at - N;
            dh += (int64_t)P->la[tt] * (1 - 2 * A + 2 * (int64_t)P->a0[tt]);
            dh += (int64_t)P->lp[tt] * dP * (dP + 2 * (Pm - P->p0[tt]));
        }
        const bool accept = dh <= 0 || (dh < a.nthr && w[2] < a.thr[dh]);
        if (!accept) continue;

        lat[ti] = vs;  // kernels2d.cu:548
        const int pos[3] = {tx, ty, tz};
        if (cs) {
            const int64_t delta = (int64_t)(N - 2 * as) * 4294967296LL + 1;
            atomic_add_i64(a.W + cs, delta);
            atomic_add_i64(a.Dcur + cs, delta);
#pragma unroll
            for (int ax = 0; ax < D; ax++) {
                const int v = pos[ax] - a.anchor[3 * cs + ax];
                atomic_add_i64(a.csum + 3 * cs + ax, (int64_t)(((v + (Lm + 1) / 2) & Lm) - (Lm + 1) / 2));
            }
        }
        if (ct) {
            const int64_t delta = (int64_t)(2 * at - N) * 4294967296LL - 1;
            atomic_add_i64(a.W + ct, delta);
            atomic_add_i64(a.Dcur + ct, delta);
#pragma unroll
            for (int ax = 0; ax < D; ax++) {
                const int v = pos[ax] - a.anchor[3 * ct + ax];
                atomic_add_i64(a.csum + 3 * ct + ax, -(int64_t)(((v + (Lm + 1) / 2) & Lm) - (Lm + 1) / 2));
            }
        }
    }
}

// Area and perimeter recount (kernels2d.cu:5-41): one thread per site.
template <int D>
__global__ __launch_bounds__(256) void k_tally_area_perim(const uint32_t* __restrict__ lat, int lgL,
                                                          int64_t* tab, uint32_t* first) {
    const uint32_t nsites = 1u << (D * lgL);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nsites) return;
    const uint32_t c = lat[i] & kSpinMask;
    if (!c) return;
    const int Lm = (1 << lgL) - 1;
    const int x = (int)(i & Lm), y = (int)((i >> lgL) & Lm), z = D == 3 ? (int)(i >> (2 * lgL)) : 0;
    constexpr int NB = D == 3 ? 27 : 9;
    int cnt = 0;
#pragma unroll
    for (int k = 0; k < NB; k++) {
        int ox, oy, oz;
        block_offset<D>(k, ox, oy, oz);
        cnt += (lat[site_index<D>((x + ox) & Lm, (y + oy) & Lm, (z + oz) & Lm, lgL)] & kSpinMask) != c;
    }
    atomic_add_i64(tab + c, (int64_t)cnt * 4294967296LL + 1);
    atomicMin(first + c, i);
}

__global__ void k_anchor(const uint32_t* __restrict__ first, int32_t* anchor, int ncells, int lgL, int d) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c > ncells) return;
    uint32_t f = first[c];
    if (f == 0xFFFFFFFFu) f = 0;
    const uint32_t Lm = (1u << lgL) - 1u;
    anchor[3 * c + 0] = (int32_t)(f & Lm);
    anchor[3 * c + 1] = (int32_t)((f >> lgL) & Lm);
    anchor[3 * c + 2] = d == 3 ? (int32_t)(f >> (2 * lgL)) : 0;
}

template <int D>
__global__ __launch_bounds__(256) void k_tally_csum(const uint32_t* __restrict__ lat, int lgL,
                                                    const int32_t* __restrict__ anchor, int64_t* csum) {
    const uint32_t nsites = 1u << (D * lgL);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nsites) return;
    const uint32_t c = lat[i] & kSpinMask;
    if (!c) return;
    const int L = 1 << lgL, Lm = L - 1;
    const int pos[3] = {(int)(i & Lm), (int)((i >> lgL) & Lm), D == 3 ? (int)(i >> (2 * lgL)) : 0};
#pragma unroll
    for (int ax = 0; ax < D; ax++) {
        const int v = pos[ax] - anchor[3 * c + ax];
        atomic_add_i64(csum + 3 * c + ax, (int64_t)(((v + L / 2) & Lm) - L / 2));
    }
}

__global__ void k_settle(int64_t* tab, int64_t* del, int ncells) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c > ncells) return;
    tab[c] += del[c];
    del[c] = 0;
}

hipError_t launch_phase(int d, const PhaseArgs& a, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
    const int work = a.nsub > a.ncells + 1 ? a.nsub : a.ncells + 1;
    const int blocks = (work + 255) / 256;
    // hipExtLaunchKernelGGL records e0/e1 at the dispatch's start and end
    // (null events: plain launch).
    if (d == 2)
        hipExtLaunchKernelGGL(k_sweep_phase<2>, dim3(blocks), dim3(256), 0, s, e0, e1, 0, a);
    else
        hipExtLaunchKernelGGL(k_sweep_phase<3>, dim3(blocks), dim3(256), 0, s, e0, e1, 0, a);
    return hipGetLastError();
}

hipError_t launch_tally(int d, const uint32_t* lat, int lgL, int ncells, int64_t* tab,
                        uint32_t* first, int32_t* anchor, int64_t* csum, hipStream_t s) {
    const uint32_t nsites = 1u << (d * lgL);
    const int blocks = (int)((nsites + 255) / 256);
    if (d == 2)
        hipLaunchKernelGGL(k_tally_area_perim<2>, dim3(blocks), dim3(256), 0, s, lat, lgL, tab, first);
    else
        hipLaunchKernelGGL(k_tally_area_perim<3>, dim3(blocks), dim3(256), 0, s, lat, lgL, tab, first);
    hipLaunchKernelGGL(k_anchor, dim3((ncells + 1 + 255) / 256), dim3(256), 0, s, first, anchor, ncells, lgL, d);
    if (d == 2)
        hipLaunchKernelGGL(k_tally_csum<2>, dim3(blocks), dim3(256), 0, s, lat, lgL, anchor, csum);
    else
        hipLaunchKernelGGL(k_tally_csum<3>, dim3(blocks), dim3(256), 0, s, lat, lgL, anchor, csum);
    return hipGetLastError();
}

hipError_t launch_settle(int64_t* tab, int64_t* del, int ncells, hipStream_t s) {
    hipLaunchKernelGGL(k_settle, dim3((ncells + 1 + 255) / 256), dim3(256), 0, s, tab, del, ncells);
    return hipGetLastError();
}

}  // namespace gcpm
