// kernels.hip -- HIP kernels of the MI355X CPM engine (gfx950).
//
// The checkerboard copy-attempt sweep (reference: src/kernels2d.cu:44-583,
// src/kernels3d.cu:51-594) in three launch forms that compute identical
// results: one colour phase per launch, one cooperative launch per run with
// the lattice in global memory, and (2D) one cooperative launch per run with
// each block's tile resident in LDS.  Also the tally rebuild run at push
// (kernels2d.cu:5-41, kernels3d.cu:3-47) and the end-of-run settle of the
// double-buffered cell tables.
#include <hip/hip_ext.h>

#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <mutex>

#include "cpm_device.h"

namespace gcpm {

// Philox4x32-10 (Salmon et al., SC'11).  Round r uses key + r * (W0, W1).
__device__ __forceinline__ void philox4x32_10(uint32_t x0, uint32_t x1, uint32_t x2,
                                              uint32_t x3, uint32_t k0, uint32_t k1,
                                              uint32_t w[4]) {
    // Opaque to the optimiser, so the ten round keys are computed per call
    // (scalar adds) rather than hoisted out of the caller's loop into twenty
    // long-lived scalar registers.
    // (The key is the same for every lane.)
    k0 = __builtin_amdgcn_readfirstlane(k0);
    k1 = __builtin_amdgcn_readfirstlane(k1);
    asm volatile("" : "+s"(k0), "+s"(k1));
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
    // k in [0, 27): k / 3 = (11 k) >> 5 and k / 9 = (57 k) >> 9 are exact
    // there (checked for every k), and take 24-bit multiplies
    const uint32_t u = (uint32_t)k;
    const uint32_t d3 = (u * 11u) >> 5, d9 = (u * 57u) >> 9;
    ox = (int)(u - 3u * d3) - 1;
    oy = (int)(d3 - 3u * d9) - 1;
    oz = D == 3 ? (int)d9 - 1 : 0;
}

template <int D>
__device__ __forceinline__ uint32_t site_index(int x, int y, int z, int lgL) {
    return (uint32_t)x | ((uint32_t)y << lgL) | (D == 3 ? ((uint32_t)z << (2 * lgL)) : 0u);
}

__device__ __forceinline__ void atomic_add_i64(int64_t* p, int64_t v) {
    atomicAdd(reinterpret_cast<unsigned long long*>(p), (unsigned long long)v);
}

// Device-coherent loads and stores (agent-scope relaxed atomics; on gfx950
// they carry the sc1 bit and bypass the per-XCD caches).  The persistent
// kernels use them for every word that one block writes and another block
// reads in a later phase, so a phase boundary needs no L2 writeback or
// invalidate, only completion of the stores (grid_barrier).
template <bool CO, class T>
__device__ __forceinline__ T ld(const T* p) {
    if constexpr (CO)
        return __hip_atomic_load(const_cast<T*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else
        return *p;
}

template <bool CO, class T>
__device__ __forceinline__ void st(T* p, T v) {
    if constexpr (CO)
        __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else
        *p = v;
}

// Lattice held in global memory (HBM / L2 / MALL); CO selects
// device-coherent accesses (persistent kernels).
template <int D, bool CO, class T = uint32_t>
struct GlobalLat {
    using Site = T;
    T* p;
    int lgL, Lm;
    struct Pos {
        int x, y, z;
        uint32_t i;
    };
    __device__ __forceinline__ Pos at(int x, int y, int z) const {
        return {x, y, z, site_index<D>(x, y, z, lgL)};
    }
    __device__ __forceinline__ uint32_t self(const Pos& q) const { return ld<CO>(p + q.i); }
    __device__ __forceinline__ uint32_t nb(const Pos& q, int dx, int dy, int dz) const {
        return ld<CO>(p + site_index<D>((q.x + dx) & Lm, (q.y + dy) & Lm, (q.z + dz) & Lm, lgL));
    }
    __device__ __forceinline__ void put(const Pos& q, uint32_t v) const { st<CO>(p + q.i, (T)v); }
    __device__ __forceinline__ uint32_t gindex(const Pos& q) const { return q.i; }
    // The 3^D block around q (centre included), (dz, dy, dx) order with dx
    // fastest.  Byte offsets from the lattice base fit in 32 bits (at most
    // 2^30 sites), so every load is base (scalar) + one 32-bit offset.
    __device__ __forceinline__ void block(const Pos& q, uint32_t* nv) const {
        constexpr int SH = sizeof(T) == 2 ? 1 : 2;  // log2 bytes per site
        const char* base = reinterpret_cast<const char*>(p);
        const uint32_t xo[3] = {(uint32_t)((q.x - 1) & Lm) << SH, (uint32_t)q.x << SH,
                                (uint32_t)((q.x + 1) & Lm) << SH};
        const uint32_t yo[3] = {(uint32_t)((q.y - 1) & Lm) << (lgL + SH), (uint32_t)q.y << (lgL + SH),
                                (uint32_t)((q.y + 1) & Lm) << (lgL + SH)};
        uint32_t zo[3] = {0u, 0u, 0u};
        if constexpr (D == 3) {
            zo[0] = (uint32_t)((q.z - 1) & Lm) << (2 * lgL + SH);
            zo[1] = (uint32_t)q.z << (2 * lgL + SH);
            zo[2] = (uint32_t)((q.z + 1) & Lm) << (2 * lgL + SH);
        }
        constexpr int NB = D == 3 ? 27 : 9;
#pragma unroll
        for (int n = 0; n < NB; n++)
            nv[n] = ld<CO>(reinterpret_cast<const T*>(base + (xo[n % 3] | yo[(n / 3) % 3] | zo[n / 9])));
    }
};

// The global lattice of a sweep in site type T: the 32-bit lattice, or its
// 16-bit form (SweepConsts::lat16).
template <class T>
__device__ __forceinline__ T* lat_base(const SweepConsts& k) {
    if constexpr (sizeof(T) == 2)
        return k.lat16;
    else
        return k.lat;
}

// Neighbourhood loads of a lattice held in LDS.
template <int D, class Lat, class Pos>
__device__ __forceinline__ void lds_block(const Lat& lat, const Pos& q, uint32_t* nv) {
    constexpr int NB = D == 3 ? 27 : 9;
#pragma unroll
    for (int n = 0; n < NB; n++) {
        int nx, ny, nz;
        block_offset<D>(n, nx, ny, nz);
        nv[n] = lat.nb(q, nx, ny, nz);
    }
}

// 2D tile of the lattice resident in LDS with a one-site halo ring; row
// stride S = tile width + 2.  Writes go to LDS and through to global memory,
// where neighbouring tiles read them as their halo in the next phase.
struct LdsLat2 {
    using Site = uint32_t;
    uint32_t* s;
    uint32_t* g;
    int S, x0, y0, lgL;
    struct Pos {
        int li;
        uint32_t gi;
    };
    __device__ __forceinline__ Pos at(int x, int y, int) const {
        return {(y - y0 + 1) * S + (x - x0 + 1), (uint32_t)x | ((uint32_t)y << lgL)};
    }
    __device__ __forceinline__ uint32_t self(const Pos& q) const { return s[q.li]; }
    __device__ __forceinline__ uint32_t nb(const Pos& q, int dx, int dy, int) const {
        return s[q.li + dy * S + dx];
    }
    __device__ __forceinline__ void put(const Pos& q, uint32_t v) const {
        s[q.li] = v;
        st<true>(g + q.gi, v);
    }
    __device__ __forceinline__ void block(const Pos& q, uint32_t* nv) const {
        lds_block<2>(*this, q, nv);
    }
    __device__ __forceinline__ uint32_t gindex(const Pos& q) const { return q.gi; }
};

// A 3D box of the lattice staged in LDS for one colour phase, with a
// one-site halo; strides SX (x) and SXY (z).  Writes go to LDS and to global
// memory (plain stores: the next launch sees them).
struct LdsLat3 {
    using Site = uint32_t;
    uint32_t* s;
    uint32_t* g;
    int SX, SXY, x0, y0, z0, lgL;
    struct Pos {
        int li;
        uint32_t gi;
    };
    __device__ __forceinline__ Pos at(int x, int y, int z) const {
        return {(z - z0 + 1) * SXY + (y - y0 + 1) * SX + (x - x0 + 1),
                (uint32_t)x | ((uint32_t)y << lgL) | ((uint32_t)z << (2 * lgL))};
    }
    __device__ __forceinline__ uint32_t self(const Pos& q) const { return s[q.li]; }
    __device__ __forceinline__ uint32_t nb(const Pos& q, int dx, int dy, int dz) const {
        return s[q.li + dz * SXY + dy * SX + dx];
    }
    __device__ __forceinline__ void put(const Pos& q, uint32_t v) const {
        s[q.li] = v;
        g[q.gi] = v;
    }
    __device__ __forceinline__ void block(const Pos& q, uint32_t* nv) const {
        lds_block<3>(*this, q, nv);
    }
    __device__ __forceinline__ uint32_t gindex(const Pos& q) const { return q.gi; }
};

// Cell-table buffers of colour phase g (see SweepConsts).  row() reads a
// cell's phase-start row, add() commits an accepted flip's delta.
struct PhaseTables {
    const int64_t* R;   // tab[g & 1]: the phase-start snapshot
    int64_t* W;         // tab[(g + 1) & 1]: becomes the state after phase g
    int64_t* Dc;        // del[g & 1]: deltas of this phase (null: W alone takes them)
    int64_t* Dp;        // del[(g + 1) & 1]: deltas of phase g - 1, folded into W
    template <bool CO>
    __device__ __forceinline__ int64_t row(uint32_t c) const { return ld<CO>(R + c); }
    __device__ __forceinline__ void add(uint32_t c, int64_t delta) const {
        atomic_add_i64(W + c, delta);
        if (Dc) atomic_add_i64(Dc + c, delta);
    }
};

__device__ __forceinline__ PhaseTables tables_of(const SweepConsts& k, int g) {
    // selects rather than an index: a dynamic index into a kernel's private
    // copy of SweepConsts would place the copy in scratch memory
    const bool odd = g & 1;
    return {odd ? k.tab[1] : k.tab[0], odd ? k.tab[0] : k.tab[1], odd ? k.del[1] : k.del[0],
            odd ? k.del[0] : k.del[1]};
}

// One copy attempt, split in two: draw_attempt makes the Philox draw and
// issues the loads of the target's 3^D neighbourhood; finish_attempt
// evaluates dH and commits.
template <int D, class Lat>
struct Draw {
    static constexpr int NB = D == 3 ? 27 : 9;
    uint32_t w[4];
    typename Lat::Pos t;
    uint32_t nv[NB];  // the 3^D block around the target, centre included
    uint32_t vs;      // the source: the neighbour picked by w[1]
    int dir;          // the source's index in the 3^D block
    uint32_t tick;    // MCS counter (act term)
};

template <int D>
__device__ __forceinline__ uint32_t sub_block_id(const SweepConsts& k, int bx, int by, int bz) {
    const int lsb = k.lgL - k.lc;  // log2 sub-blocks per axis
    return (uint32_t)bx | ((uint32_t)by << lsb) | (D == 3 ? ((uint32_t)bz << (2 * lsb)) : 0u);
}

// The random words of attempt i of sub-block sb (x coordinate bx) in one
// colour phase: word 0 picks the position in the sub-block, word 1 the
// neighbour, word 2 the acceptance.  With one position per sub-block (lc = 0)
// an attempt needs only words 1 and 2, so the same-colour sub-blocks 2j and
// 2j + 1 along x (colour-grid column bx >> lstride) share one Philox block:
// the even one's counter, words 1 and 2 to the even sub-block, words 3 and 0
// to the odd one.
__device__ __forceinline__ void attempt_words(const SweepConsts& k, uint32_t sb, int bx,
                                              uint32_t mcs, uint32_t phase, uint32_t i,
                                              uint32_t w[4]) {
    const uint32_t odd = k.lc == 0 ? ((uint32_t)bx >> k.lstride) & 1u : 0u;
    philox4x32_10(sb - (odd << k.lstride), mcs, phase, i, k.k0, k.k1, w);
    if (odd) {
        w[1] = w[3];
        w[2] = w[0];
    }
}

// Source direction (kernels2d.cu:216-225 drawn in one step): floor(w1 N /
// 2^32) over the N neighbours in block order, skipping the centre.
// In 2D the source is selected from the nine loaded values, so all loads
// are issued in one round.  In 3D it is loaded on its own: the compiler then
// issues the target and source loads first and the other 25 only when they
// differ, which skips most of the loads (and work) in sparse lattices.
template <int D, class Lat>
__device__ __forceinline__ void draw_attempt(const Lat& lat, const SweepConsts& k, uint32_t sb,
                                             int bx, int by, int bz, uint32_t mcs,
                                             uint32_t phase, uint32_t i, Draw<D, Lat>& d) {
    const int lc = k.lc;
    const uint32_t cm = (1u << lc) - 1u;
    attempt_words(k, sb, bx, mcs, phase, i, d.w);
    d.tick = mcs;
    constexpr int N = D == 3 ? 26 : 8;
    constexpr int NB = N + 1;
    int px = (bx << lc) + (int)(d.w[0] & cm);
    int py = (by << lc) + (int)((d.w[0] >> lc) & cm);
    int pz = D == 3 ? (bz << lc) + (int)((d.w[0] >> (2 * lc)) & cm) : 0;
    int dir = (int)(((uint64_t)d.w[1] * (uint64_t)N) >> 32);
    dir += dir >= N / 2;
    if (k.roles) {
        // non-shared roles (kernels2d.cu:227-243): the drawn position is the
        // source and the target its neighbour; the source is then the
        // target's mirrored neighbour
        int ox, oy, oz;
        block_offset<D>(dir, ox, oy, oz);
        const int Lm = (1 << k.lgL) - 1;
        px = (px + ox) & Lm;
        py = (py + oy) & Lm;
        pz = D == 3 ? (pz + oz) & Lm : 0;
        dir = NB - 1 - dir;
    }
    d.t = lat.at(px, py, pz);
    lat.block(d.t, d.nv);
    d.dir = dir;
    if constexpr (D == 2) {
        uint32_t vs = 0;
#pragma unroll
        for (int n = 0; n < 9; n++) vs = n == dir ? d.nv[n] : vs;
        d.vs = vs;
    } else {
        int sx, sy, sz;
        block_offset<D>(dir, sx, sy, sz);
        d.vs = lat.nb(d.t, sx, sy, sz);
    }
}

// The attempt of kernels2d.cu:197-570 / kernels3d.cu:165-583 in the
// shared-mode roles (target drawn in the sub-block, source a uniformly drawn
// neighbour), with the deterministic RNG, snapshot tables and integer
// acceptance described in SweepConsts.
// cell_sync (kernels.cu:123-141 restated deterministically, u = 1): in the
// claim pass every accepted proposal raises claims[c] to
// (tag << 32 | ~sb) for its non-medium cells, so the lowest sub-block of the
// phase wins; the commit pass recomputes the attempt and commits only if it
// holds both claims.
struct CellSync {
    int pass;  // 0 off; cell_sync: 1 claim, 2 commit; split phase: 3 decide, 4 apply
    uint64_t* claims;
    uint32_t tag;
    uint8_t* decide;  // split phase: decisions by phase-local sub-block index
    int q;            // this attempt's sub-block index in the phase
};

// Adhesion over the neighbourhood (kernels2d.cu:359-402,477):
//   sum_n [v_n != v_s] J(t_s, t_n) - sum_n [v_n != v_t] JS(t_t, t_n).
// v_n == v_s implies t_n == t_s, so with c_t the number of neighbours of
// type t this is sum_t (J(t_s, t) - JS(t_t, t)) c_t - a_s J(t_s, t_s) +
// a_t JS(t_t, t_t), where a_s, a_t count the neighbours equal to source and
// target.  For up to six types the c_t are packed as 5-bit fields of one
// word (at most 26 neighbours), so the neighbour loop does no table lookups.
constexpr int kHistTypes = 6;

// Parameters staged in LDS (any number of types).
struct LdsParamView {
    static constexpr bool kAlwaysHist = false;
    const int32_t* w;
    int nt, n2;
    __device__ __forceinline__ bool hist() const { return nt <= kHistTypes; }
    __device__ __forceinline__ int32_t J(int a, int b) const { return w[a * nt + b]; }
    __device__ __forceinline__ int32_t JS(int a, int b) const { return w[n2 + a * nt + b]; }
    __device__ __forceinline__ int32_t la(int a) const { return w[2 * n2 + a]; }
    __device__ __forceinline__ int32_t a0(int a) const { return w[2 * n2 + nt + a]; }
    __device__ __forceinline__ int32_t lp(int a) const { return w[2 * n2 + 2 * nt + a]; }
    __device__ __forceinline__ int32_t p0(int a) const { return w[2 * n2 + 3 * nt + a]; }
    __device__ __forceinline__ int32_t adhesion(int ts, int tt, uint32_t hist, int as,
                                                int at) const {
        int32_t adh = 0;
        for (int t = 0; t < nt; t++)
            adh += (J(ts, t) - JS(tt, t)) * (int32_t)((hist >> (5 * t)) & 31u);
        return adh + at * JS(tt, tt) - as * J(ts, ts);
    }
};

__device__ __forceinline__ int32_t pick4(int i, int32_t v0, int32_t v1, int32_t v2, int32_t v3) {
    return i == 0 ? v0 : i == 1 ? v1 : i == 2 ? v2 : v3;
}

// At most four types: the padded block in registers (uniform loads, the
// same for every lane), each lane's entries picked by selects.
struct SmallParams {
    static constexpr bool kAlwaysHist = true;
    static constexpr int S = kSmallTypes;
    int32_t v[kSmallWords];
    __device__ __forceinline__ void load(const DevParams* prm) {
#pragma unroll
        for (int i = 0; i < kSmallWords; i++) v[i] = prm->s4[i];
    }
    __device__ __forceinline__ bool hist() const { return true; }
    __device__ __forceinline__ int32_t f(int field, int a) const {
        const int o = 2 * S * S + field * S;
        return pick4(a, v[o], v[o + 1], v[o + 2], v[o + 3]);
    }
    __device__ __forceinline__ int32_t la(int a) const { return f(0, a); }
    __device__ __forceinline__ int32_t a0(int a) const { return f(1, a); }
    __device__ __forceinline__ int32_t lp(int a) const { return f(2, a); }
    __device__ __forceinline__ int32_t p0(int a) const { return f(3, a); }
    __device__ __forceinline__ int32_t J(int, int) const { return 0; }  // unused
    __device__ __forceinline__ int32_t JS(int, int) const { return 0; }
    __device__ __forceinline__ int32_t adhesion(int ts, int tt, uint32_t hist, int as,
                                                int at) const {
        int32_t adh = 0;
#pragma unroll
        for (int t = 0; t < S; t++) {
            const int32_t js = pick4(ts, v[t], v[S + t], v[2 * S + t], v[3 * S + t]);
            const int32_t jt = pick4(tt, v[S * S + t], v[S * S + S + t], v[S * S + 2 * S + t],
                                     v[S * S + 3 * S + t]);
            adh += (js - jt) * (int32_t)((hist >> (5 * t)) & 31u);
        }
        const int32_t jss = pick4(ts, v[0], v[S + 1], v[2 * S + 2], v[3 * S + 3]);
        const int32_t jtt = pick4(tt, v[S * S], v[S * S + S + 1], v[S * S + 2 * S + 2],
                                  v[S * S + 3 * S + 3]);
        return adh + at * jtt - as * jss;
    }
};

// Chemotaxis (kernels2d.cu:325-334, kernels3d.cu:314-325) in fixed point:
// -lambda (d . f(s)) with d = (t - s) / |t - s| = -o / |o|, o = s - t the
// source's block offset, is +lambda (o . f(s)) / sqrt(k), k = |o|^2.  lambda
// is that of the source type, or of the target type when the source is
// medium.  With m = llround(lambda scale 2^8 / sqrt(k)) and the field held
// as round(f 2^16), the term is (m g + 2^23) >> 24, g = sum_i o_i f_i(s).
template <int D, class Lat>
__device__ __forceinline__ int64_t chem_energy(const SweepConsts& k, const Lat& lat,
                                               const typename Lat::Pos& t, int dir, int ts,
                                               int tt) {
    int ox, oy, oz;
    block_offset<D>(dir, ox, oy, oz);
    const int tl = ts ? ts : tt;
    const int kk = (ox != 0) + (oy != 0) + (oz != 0);
    const int64_t m = k.prm->chem_m[tl][kk];
    if (m == 0) return 0;
    const int lgL = k.lgL, Lm = (1 << lgL) - 1;
    const uint32_t gi = lat.gindex(t);
    const int x = (int)(gi & Lm), y = (int)((gi >> lgL) & Lm), z = D == 3 ? (int)(gi >> (2 * lgL)) : 0;
    const uint32_t src = site_index<D>((x + ox) & Lm, (y + oy) & Lm, (z + oz) & Lm, lgL);
    const uint32_t n = 1u << (D * lgL);
    int64_t g = (int64_t)ox * k.field[src] + (int64_t)oy * k.field[n + src];
    if constexpr (D == 3) g += (int64_t)oz * k.field[2 * n + src];
    return (m * g + ((int64_t)1 << 23)) >> 24;
}

// Act (kernels2d.cu:312-343,386-397,417-473; kernels3d.cu likewise) in
// fixed point.  A site's act value is min(max(actMax + act - tick, 0),
// actMax); G(site) is the geometric mean of the values of the site and its
// neighbours holding the same cell value (0 for medium or when a value is 0),
// 2^(mean log2) from integer tables, in 2^-16 units.  The term is
// -lambda/actMax (G_source - G_target), lambda and actMax of the target type
// (of the source type when the target is medium): (-mA dG + 2^31) >> 32.
__device__ __forceinline__ int64_t act_geo_mean(const SweepConsts& k, int64_t sum_lg, int count) {
    const int64_t mean = sum_lg / count;
    const int64_t I = mean >> 24, f = mean & 0xFFFFFF;
    const int64_t j = f >> 16, r = f & 0xFFFF;
    const int64_t X = (r * k.act_ln2q) >> 24;
    const int64_t P = ((int64_t)1 << 30) + X + ((X * X) >> 31);
    const int64_t M = (k.act_tab[j] * P) >> 30;
    return (M << I) >> 14;
}

// G of the site at global index gi holding value v, whose 3^D block values
// are nv (the centre is the site itself).
template <int D, bool CO>
__device__ __forceinline__ int64_t act_site_mean(const SweepConsts& k, uint32_t gi, uint32_t v,
                                                 const uint32_t* nv, uint32_t tick) {
    if ((v & k.idmask) == 0) return 0;
    constexpr int NB = D == 3 ? 27 : 9;
    const int lgL = k.lgL, Lm = (1 << lgL) - 1;
    const int x = (int)(gi & Lm), y = (int)((gi >> lgL) & Lm), z = D == 3 ? (int)(gi >> (2 * lgL)) : 0;
    const int64_t am = k.prm->act_max[v >> k.idbits];
    int64_t sum = 0;
    int count = 0;
    bool zero = false;
#pragma unroll
    for (int n = 0; n < NB; n++) {
        if (nv[n] != v) continue;  // the centre always matches
        int ox, oy, oz;
        block_offset<D>(n, ox, oy, oz);
        const uint32_t si = site_index<D>((x + ox) & Lm, (y + oy) & Lm, (z + oz) & Lm, lgL);
        int64_t a = am + (int64_t)ld<CO>(k.act_lat + si) - (int64_t)tick;
        a = a > am ? am : a;
        if (a <= 0)
            zero = true;
        else
            sum += k.act_tab[256 + a];
        count++;
    }
    return zero ? 0 : act_geo_mean(k, sum, count);
}

template <int D, bool CO, class Lat>
__device__ __forceinline__ int64_t act_energy(const SweepConsts& k, const Lat& lat,
                                              const typename Lat::Pos& t, const uint32_t* nvt,
                                              int dir, uint32_t vs, uint32_t vt, uint32_t tick) {
    const int tl = (vt >> k.idbits) ? (int)(vt >> k.idbits) : (int)(vs >> k.idbits);
    const int64_t m = k.prm->act_m[tl];
    if (m == 0) return 0;
    constexpr int NB = D == 3 ? 27 : 9;
    const int lgL = k.lgL, Lm = (1 << lgL) - 1;
    const uint32_t gt = lat.gindex(t);
    const int x = (int)(gt & Lm), y = (int)((gt >> lgL) & Lm), z = D == 3 ? (int)(gt >> (2 * lgL)) : 0;
    int ox, oy, oz;
    block_offset<D>(dir, ox, oy, oz);
    const int sx = (x + ox) & Lm, sy = (y + oy) & Lm, sz = (z + oz) & Lm;
    const uint32_t gs = site_index<D>(sx, sy, sz, lgL);
    // the source's block from global memory (up to two sites from the
    // target: untouched in this phase when positions_per_checkerboard >= 2)
    int64_t g_s = 0;
    if (vs & k.idmask) {
        uint32_t nvs[NB];
#pragma unroll
        for (int n = 0; n < NB; n++) {
            int px, py, pz;
            block_offset<D>(n, px, py, pz);
            nvs[n] = ld<CO>(lat_base<typename Lat::Site>(k) +
                             site_index<D>((sx + px) & Lm, (sy + py) & Lm, (sz + pz) & Lm, lgL));
        }
        g_s = act_site_mean<D, CO>(k, gs, vs, nvs, tick);
    }
    const int64_t g_t = act_site_mean<D, CO>(k, gt, vt, nvt, tick);
    return (-m * (g_s - g_t) + ((int64_t)1 << 31)) >> 32;
}

// Persistence (kernels3d.cu:327-344, 3D): -lambda (d . p) with p the
// preferred direction of the source cell (of the target cell when the source
// is medium) and lambda of the source type (target type likewise); fixed
// point as for chemotaxis, p held as llround(p 2^16).
template <int D>
__device__ __forceinline__ int64_t pers_energy(const SweepConsts& k, int dir, uint32_t vs,
                                               uint32_t vt) {
    const uint32_t cs = vs & k.idmask, ct = vt & k.idmask;
    const int tl = cs ? (int)(vs >> k.idbits) : (int)(vt >> k.idbits);
    const uint32_t cell = cs ? cs : ct;
    int ox, oy, oz;
    block_offset<D>(dir, ox, oy, oz);
    const int kk = (ox != 0) + (oy != 0) + (oz != 0);
    const int64_t m = k.prm->pers_m[tl][kk];
    if (m == 0 || cell == 0) return 0;
    const int32_t* p = k.pref + 3 * cell;
    const int64_t g = (int64_t)ox * p[0] + (int64_t)oy * p[1] + (int64_t)oz * p[2];
    return (m * g + ((int64_t)1 << 23)) >> 24;
}

// Adhesion over the 3^D block nv (centre included; kernels2d.cu:359-402,477)
// of a copy of source value vs (type ts) onto target value vt (type tt), and
// the numbers of neighbours equal to the source (as) and to the target (at).
template <int D, class PV>
__device__ __forceinline__ int32_t adhesion_energy(const SweepConsts& k, const PV& P, const uint32_t* nv,
                                                   uint32_t vs, uint32_t vt, int ts, int tt, int& as, int& at) {
    constexpr int NB = D == 3 ? 27 : 9;
    constexpr int CENTRE = NB / 2;
    int32_t adh = 0;
    as = 0;
    at = 0;
    if (PV::kAlwaysHist || P.hist()) {
        uint32_t hist = 0;
#pragma unroll
        for (int n = 0; n < NB; n++) {
            if (n == CENTRE) continue;
            const uint32_t vn = nv[n];
            hist += 1u << ((vn >> k.idbits) * 5u);
            as += vn == vs;
            at += vn == vt;
        }
        adh = P.adhesion(ts, tt, hist, as, at);
    } else {
#pragma unroll
        for (int n = 0; n < NB; n++) {
            if (n == CENTRE) continue;
            const uint32_t vn = nv[n];
            const int tn = (int)(vn >> k.idbits);
            const int32_t js = P.J(ts, tn), jt = P.JS(tt, tn);
            const bool eqs = vn == vs, eqt = vn == vt;
            adh += (eqs ? 0 : js) - (eqt ? 0 : jt);
            as += eqs;
            at += eqt;
        }
    }
    return adh;
}

// Perimeter and area terms (kernels2d.cu:479-520) of a copy from
// cell cs (type ts, phase-start row row_s) onto cell ct (tt, row_t); as and
// at as adhesion_energy gives them.  Medium (cell 0) has no terms.
template <int D, class PV>
__device__ __forceinline__ int64_t row_energy(const PV& P, uint32_t cs, uint32_t ct, int ts, int tt,
                                              int64_t row_s, int64_t row_t, int as, int at) {
    constexpr int N = D == 3 ? 26 : 8;
    int64_t dh = 0;
    if (cs) {
        const int64_t A = (int64_t)(uint32_t)row_s, Pm = row_s >> 32;
        const int64_t dP = N - 2 * as;
        dh += (int64_t)P.la(ts) * (1 + 2 * A - 2 * (int64_t)P.a0(ts));
        dh += (int64_t)P.lp(ts) * dP * (dP + 2 * (Pm - P.p0(ts)));
    }
    if (ct) {
        const int64_t A = (int64_t)(uint32_t)row_t, Pm = row_t >> 32;
        const int64_t dP = 2 * at - N;
        dh += (int64_t)P.la(tt) * (1 - 2 * A + 2 * (int64_t)P.a0(tt));
        dh += (int64_t)P.lp(tt) * dP * (dP + 2 * (Pm - P.p0(tt)));
    }
    return dh;
}

// Attempt counters (cpm_set_counters; the reference allocates analytics[3]
// for evaluated / accepted / committed and leaves its increments commented
// out, kernels2d.cu:292,526,531): the lanes that reach one point together
// add their number with one atomic per wave.
__device__ __forceinline__ void count_attempts(unsigned long long* c, int slot) {
    const unsigned long long m = __ballot(1);
    if (__lane_id() == (unsigned int)(__ffsll(m) - 1)) atomicAdd(c + slot, (unsigned long long)__popcll(m));
}

// XT = false leaves out the chemotaxis, act and persistence terms (the
// caller launches it only when none is on): fewer registers in 3D.
template <int D, bool CO, bool XT = true, class Lat, class PV, class TB>
__device__ __forceinline__ void finish_attempt(const Lat& lat, const SweepConsts& k, const PV& P,
                                               const uint32_t* __restrict__ thr,
                                               const TB& T, const CellSync& cy,
                                               uint32_t sb, const Draw<D, Lat>& d) {
    constexpr int N = D == 3 ? 26 : 8;
    constexpr int NB = N + 1;
    constexpr int CENTRE = NB / 2;
    const uint32_t vt = d.nv[CENTRE];
    uint32_t vs = d.vs;
    if (vs == vt) return;  // kernels2d.cu:276
    // counted once per attempt: in its only pass, or in the commit / apply
    // pass of a claimed or split phase
    unsigned long long* const ctr = (cy.pass == 0 || cy.pass == 2 || cy.pass == 4 || cy.pass == 6) ? k.counters : nullptr;
    if (ctr) count_attempts(ctr, 0);
    if (k.fixed_mask) {    // fixed types (kernels2d.cu:281-290)
        if ((k.fixed_mask >> (vt >> k.idbits)) & 1u) return;  // a fixed target stays
        if ((k.fixed_mask >> (vs >> k.idbits)) & 1u) {        // a fixed source copies medium
            vs = 0;
            // Onto a medium target this is a medium -> medium attempt, which
            // the reference evaluates and, if accepted, stamps act[target]
            // (kernels2d.cu:532-534); nothing else changes, so without the
            // act term it is skipped.
            if (vs == vt && !(XT && k.act)) return;
        }
    }

    if (ctr) count_attempts(ctr, 1);  // kernels2d.cu:292
    const int ts = (int)(vs >> k.idbits), tt = (int)(vt >> k.idbits);
    const uint32_t cs = vs & k.idmask, ct = vt & k.idmask;
    // Issue the snapshot-table loads first: their latency overlaps the
    // neighbourhood sums.
    const int64_t row_s = cs ? T.template row<CO>(cs) : 0;
    const int64_t row_t = ct ? T.template row<CO>(ct) : 0;
    int as, at;
    int64_t dh = adhesion_energy<D>(k, P, d.nv, vs, vt, ts, tt, as, at);
    dh += row_energy<D>(P, cs, ct, ts, tt, row_s, row_t, as, at);
    bool accept;
    if (cy.pass == 4 || cy.pass == 6) {
        // split phase, apply: the decision made against the phase-start
        // lattice (the act term reads sites that other attempts now write)
        accept = cy.decide[cy.q] != 0;
    } else {
        if (XT && k.chem) dh += chem_energy<D>(k, lat, d.t, d.dir, ts, tt);
        if (XT && k.act) dh += act_energy<D, CO>(k, lat, d.t, d.nv, d.dir, vs, vt, d.tick);
        if (XT && D == 3 && k.pers) dh += pers_energy<D>(k, d.dir, vs, vt);
        accept = dh <= 0 || (dh < k.nthr && d.w[2] < thr[dh]);
    }
    if (cy.pass == 3 || cy.pass == 5) {  // split phase, decide: record (pass 5 also claims)
        if (accept) cy.decide[cy.q] = 1;
        if (cy.pass == 3) return;
    }
    if (!accept) return;
    if (ctr) count_attempts(ctr, 2);  // kernels2d.cu:526
    if (cy.pass == 1 || cy.pass == 2 || cy.pass == 5 || cy.pass == 6) {
        const uint64_t key = ((uint64_t)cy.tag << 32) | (uint64_t)(0xFFFFFFFFu - sb);
        if (cy.pass == 1 || cy.pass == 5) {
            if (cs) atomicMax(reinterpret_cast<unsigned long long*>(cy.claims + cs), key);
            if (ct) atomicMax(reinterpret_cast<unsigned long long*>(cy.claims + ct), key);
            return;
        }
        if ((cs && cy.claims[cs] != key) || (ct && cy.claims[ct] != key)) return;
    }

    if (ctr) count_attempts(ctr, 3);  // kernels2d.cu:531
    lat.put(d.t, vs);  // kernels2d.cu:548
    if (XT && k.act) st<CO>(k.act_lat + lat.gindex(d.t), d.tick);  // kernels2d.cu:532-534
    // Area and perimeter deltas (kernels2d.cu:552-569); centroid sums are
    // rebuilt from the lattice when read.
    if (cs) T.add(cs, (int64_t)(N - 2 * as) * 4294967296LL + 1);
    if (ct) T.add(ct, (int64_t)(2 * at - N) * 4294967296LL - 1);
}

// MINSTD (random.cu:5-12): 16807 x mod (2^31 - 1), in double as there.
__device__ __forceinline__ int32_t minstd_next(int32_t x) {
    const double temp = (double)x * 16807.0;
    return (int32_t)(temp - 2147483647.0 * floor(temp * (1.0 / 2147483647.0)));
}

// At most this many direction triples per attempt (a stream stuck at one
// value would otherwise never end; a real one needs more than a few with
// probability 9^-n).
constexpr int kRefTries = 256;

// The attempts of sub-block (bx, by, bz) drawn from its region's reference
// stream, in the reference's order (kernels2d.cu:199-243,276-290,522;
// kernels3d.cu:160-215): d position draws mod c; direction triples
// (each u mod 3 - 1) until one passes the reference's rejection test (all
// coordinates equal; in 3D without shared roles it compares z with y,
// kernels3d.cu:208); then, when the sites differ and the target type is not
// fixed, the acceptance draw.  A claim pass (cell_sync) leaves the state
// for its commit pass to redo.
template <int D, bool CO, class Lat, class PV>
__device__ __forceinline__ void ref_attempts(const Lat& lat, const SweepConsts& k, const PV& P,
                                             const uint32_t* __restrict__ thr, int bx, int by,
                                             int bz, uint32_t mcs, const PhaseTables& T,
                                             const CellSync& cy, int natt) {
    constexpr int NB = D == 3 ? 27 : 9;
    const uint32_t sb = sub_block_id<D>(k, bx, by, bz);
    if (natt < 0) natt = k.u;
    // the reference thread whose region holds the sub-block (with lg = 0 the
    // region itself; see ref_region)
    auto thread_of = [&](int b) {
        const int r = b >> k.lk;
        return (uint32_t)(((r >> (k.lg + k.ld)) << k.ld) | (r & ((1 << k.ld) - 1)));
    };
    const uint32_t rid = thread_of(bx) | (thread_of(by) << k.lr) | (D == 3 ? thread_of(bz) << (2 * k.lr) : 0u);
    int32_t st = k.rstate[rid];
    const int c = 1 << k.lc, Lm = (1 << k.lgL) - 1;
    for (int i = 0; i < natt; i++) {
        st = minstd_next(st);
        const int px = (bx << k.lc) + st % c;
        st = minstd_next(st);
        const int py = (by << k.lc) + st % c;
        int pz = 0;
        if (D == 3) {
            st = minstd_next(st);
            pz = (bz << k.lc) + st % c;
        }
        int dx = 0, dy = 0, dz = 0, qx = px, qy = py, qz = pz;
        for (int tries = 0; tries < kRefTries; tries++) {
            st = minstd_next(st);
            dx = st % 3 - 1;
            st = minstd_next(st);
            dy = st % 3 - 1;
            if (D == 3) {
                st = minstd_next(st);
                dz = st % 3 - 1;
            }
            qx = (px + dx) & Lm;
            qy = (py + dy) & Lm;
            qz = D == 3 ? (pz + dz) & Lm : 0;
            const bool rej = D == 2  ? (qx == px && qy == py)
                             : k.roles ? (qx == px && qy == py && qz == py)
                                       : (qx == px && qy == py && qz == pz);
            if (!rej) break;
        }
        // shared roles: the drawn site is the target; otherwise the source
        const int tx = k.roles ? qx : px, ty = k.roles ? qy : py, tz = k.roles ? qz : pz;
        const int ox = k.roles ? -dx : dx, oy = k.roles ? -dy : dy, oz = k.roles ? -dz : dz;
        Draw<D, Lat> d;
        d.t = lat.at(tx, ty, tz);
        lat.block(d.t, d.nv);
        d.dir = D == 3 ? (oz + 1) * 9 + (oy + 1) * 3 + (ox + 1) : (oy + 1) * 3 + (ox + 1);
        uint32_t vs = 0;
#pragma unroll
        for (int n = 0; n < NB; n++) vs = n == d.dir ? d.nv[n] : vs;
        d.vs = vs;
        const uint32_t vt = d.nv[NB / 2];
        d.w[0] = d.w[1] = d.w[2] = d.w[3] = 0u;
        if (vs != vt && !((k.fixed_mask >> (vt >> k.idbits)) & 1u)) {
            st = minstd_next(st);
            d.w[2] = (uint32_t)st;
        }
        d.tick = mcs;
        finish_attempt<D, CO>(lat, k, P, thr, T, cy, sb, d);
    }
    if (cy.pass != 1 && cy.pass != 3 && cy.pass != 5) k.rstate[rid] = st;
}

// The u copy attempts of sub-block (bx, by, bz) in one colour phase, in
// order.
template <int D, bool CO, bool XT = true, class Lat, class PV>
__device__ __forceinline__ void sub_block_attempts(const Lat& lat, const SweepConsts& k,
                                                   const PV& P, const uint32_t* __restrict__ thr,
                                                   int bx, int by, int bz, uint32_t mcs,
                                                   uint32_t phase, const PhaseTables& T,
                                                   const CellSync cy = CellSync{0, nullptr, 0},
                                                   int att0 = 0, int natt = -1) {
    const uint32_t sb = sub_block_id<D>(k, bx, by, bz);
    if (natt < 0) natt = k.u;
    if (natt == 1) {
        // One attempt, straight-line: with no loop the compiler keeps no
        // loop-invariant values alive (they would spill scalar registers).
        Draw<D, Lat> d;
        draw_attempt<D>(lat, k, sb, bx, by, bz, mcs, phase, (uint32_t)att0, d);
        finish_attempt<D, CO, XT>(lat, k, P, thr, T, cy, sb, d);
        return;
    }
    for (int i = att0; i < att0 + natt; i++) {
        Draw<D, Lat> d;
        draw_attempt<D>(lat, k, sb, bx, by, bz, mcs, phase, (uint32_t)i, d);
        finish_attempt<D, CO, XT>(lat, k, P, thr, T, cy, sb, d);
    }
}

// Parameters and (when short) the acceptance table staged in LDS.
constexpr int kThrLds = 1024;
struct LdsParams {
    int32_t p[kParamWords];
    uint32_t thr[kThrLds];
};

// All global loads are issued before any LDS store.
__device__ __forceinline__ const uint32_t* stage_params(LdsParams& sp, const SweepConsts& k,
                                                        LdsParamView* view) {
    constexpr int RP = (kParamWords + kBlock - 1) / kBlock;
    constexpr int RT = kThrLds / kBlock;
    const int npar = param_words(k.nt);
    const bool fits = k.nthr <= kThrLds;
    const int32_t* src = k.prm->w;
    int32_t pv[RP];
    uint32_t tv[RT];
#pragma unroll
    for (int j = 0; j < RP; j++) {
        const int w = threadIdx.x + j * kBlock;
        pv[j] = w < npar ? src[w] : 0;
    }
#pragma unroll
    for (int j = 0; j < RT; j++) {
        const int w = threadIdx.x + j * kBlock;
        tv[j] = fits && w < k.nthr ? k.thr[w] : 0u;
    }
#pragma unroll
    for (int j = 0; j < RP; j++) {
        const int w = threadIdx.x + j * kBlock;
        if (w < npar) sp.p[w] = pv[j];
    }
#pragma unroll
    for (int j = 0; j < RT; j++) {
        const int w = threadIdx.x + j * kBlock;
        if (fits && w < k.nthr) sp.thr[w] = tv[j];
    }
    __syncthreads();
    *view = LdsParamView{sp.p, k.nt, k.nt * k.nt};
    return fits ? sp.thr : k.thr;
}

// Fold of phase g: add the previous phase's deltas to W and clear them.  W
// is never read in a phase and the fold commutes with the flips' own adds,
// so it may run at any time during the phase.
template <bool CO>
__device__ __forceinline__ void fold_tables(const PhaseTables& T, int ncells, int gtid, int gsz) {
    for (int ci = gtid; ci <= ncells; ci += gsz) {
        const int64_t dv = ld<CO>(T.Dp + ci);
        if (dv != 0) {
            atomic_add_i64(T.W + ci, dv);
            st<CO>(T.Dp + ci, (int64_t)0);
        }
    }
}

// Colour of phase index col: cbX outermost, then cbY (, cbZ) (kernels2d.cu:179-180).
template <int D>
__device__ __forceinline__ void colour_of(int col, int& cx, int& cy, int& cz) {
    cx = D == 3 ? (col >> 2) & 1 : (col >> 1) & 1;
    cy = D == 3 ? (col >> 1) & 1 : col & 1;
    cz = D == 3 ? col & 1 : 0;
}

// The reference's shared mode with positions_per_thread > 2c
// (kernels2d.cu:72-86,92-106,136-146): blocks of 2^ld threads per axis,
// each thread owning 2^lg regions of 2c sites per axis, visited one block
// slice g at a time, so thread j of a phase in slice g works in region
// ((j >> ld) << (lg + ld)) | (g << ld) | (j mod 2^ld).
__device__ __forceinline__ int ref_region(const SweepConsts& k, int j, int g) {
    return ((j >> k.ld) << (k.lg + k.ld)) | (g << k.ld) | (j & ((1 << k.ld) - 1));
}

// One colour phase.  Thread q owns same-colour sub-block q and makes its u
// copy attempts in order; sub-blocks of one colour are two sub-blocks apart,
// so no attempt reads a site another thread writes in this phase.  The
// first ncells + 1 threads also fold the previous phase's deltas (their
// loads are issued before the attempts, the adds made after them).
template <int D, class ST, bool REF = false, class PV>
__device__ __forceinline__ void sweep_phase_body(const PhaseArgs& a, const PV& P,
                                                 const uint32_t* __restrict__ thr) {
    const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const int gsz = gridDim.x * blockDim.x;
    const SweepConsts& k = a.k;
    const PhaseTables T = tables_of(k, a.g);
    const bool fold = a.sync_pass != 1 && a.sync_pass != 3 && a.sync_pass != 5;
    const int64_t dv = fold && gtid <= k.ncells ? T.Dp[gtid] : 0;
    if (gtid < a.nsub) {
        if (a.sync_pass == 3 || a.sync_pass == 5) a.decide[gtid] = 0;
        const int hm = (1 << a.lh) - 1;
        int jx = gtid & hm, jy = (gtid >> a.lh) & hm, jz = D == 3 ? gtid >> (2 * a.lh) : 0;
        if (REF && k.lg) {
            jx = ref_region(k, jx, a.gsx);
            jy = ref_region(k, jy, a.gsy);
            if (D == 3) jz = ref_region(k, jz, a.gsz);
        }
        const int bx = (jx << a.lstride) + a.cx;
        const int by = (jy << a.lstride) + a.cy;
        const int bz = D == 3 ? (jz << a.lstride) + a.cz : 0;
        const GlobalLat<D, false, ST> lat{lat_base<ST>(k), k.lgL, (1 << k.lgL) - 1};
        const CellSync cy{a.sync_pass, a.claims, a.sync_tag, a.decide, gtid};
        if constexpr (REF)
            ref_attempts<D, false>(lat, k, P, thr, bx, by, bz, a.mcs, T, cy, a.natt);
        else
            sub_block_attempts<D, false>(lat, k, P, thr, bx, by, bz, a.mcs, a.phase, T, cy, a.att0,
                                         a.natt);
    }
    if (fold) {
        if (dv != 0) {
            atomic_add_i64(T.W + gtid, dv);
            T.Dp[gtid] = 0;
        }
        for (int ci = gtid + gsz; ci <= k.ncells; ci += gsz) {
            const int64_t v = T.Dp[ci];
            if (v != 0) {
                atomic_add_i64(T.W + ci, v);
                T.Dp[ci] = 0;
            }
        }
    }
}

// At most four types: parameters from uniform (scalar) loads, thresholds
// read from global memory; no staging and no block barrier.
template <int D, class ST>
__global__ __launch_bounds__(kBlock) void k_sweep_phase(PhaseArgs a) {
    SmallParams P;
    P.load(a.k.prm);
    sweep_phase_body<D, ST>(a, P, a.k.thr);
}

// One attempt per sub-block (u = 1), with the attempts compacted: every
// thread draws its attempt and loads target and source (in 2D the whole
// neighbourhood); only the attempts whose source and target differ (most do
// not: cell interiors and medium) are queued in LDS, and the block's threads
// then finish the queued attempts densely.  Parameters and thresholds are
// staged in LDS while the draws' loads are in flight.  The result does not
// depend on which thread finishes an attempt.
template <int N>
struct QueueExtra {
    uint32_t x[N];
};
template <>
struct QueueExtra<0> {};

// 16 bytes per entry with 16-bit sites: the target value rides in aux
// beside the source index; 32-bit sites carry it in x[0]; FULL entries carry
// the whole 3^D block in x.  (The sub-block id is recomputed when needed.)
template <int D, bool FULL, class ST>
struct QueuedAttempt : QueueExtra<FULL ? (D == 2 ? 9 : 27) : (sizeof(ST) == 2 ? 0 : 1)> {
    static constexpr int NQ = FULL ? (D == 2 ? 9 : 27) : 0;  // neighbourhood carried
    static constexpr bool PACK = !FULL && sizeof(ST) == 2;
    uint32_t site;  // target site index
    uint32_t vs;    // source value
    uint32_t w2;    // acceptance draw
    uint32_t aux;   // source index in the 3^D block (| target value << 8 when PACK)
    __device__ __forceinline__ void set_vt(uint32_t vt) {
        if constexpr (PACK)
            aux |= vt << 8;
        else if constexpr (!FULL)
            this->x[0] = vt;
    }
    __device__ __forceinline__ int dir() const { return PACK ? (int)(aux & 0xFFu) : (int)aux; }
    __device__ __forceinline__ uint32_t vt() const {
        if constexpr (PACK)
            return aux >> 8;
        else if constexpr (!FULL)
            return this->x[0];
        else
            return this->x[NQ / 2];
    }
};

// One colour phase of one lattice as seen by a block group (the whole grid,
// or the blocks of one replica in a batch).
struct CompactPhase {
    int nsub, lh, lstride, cx, cy, cz, g;
    uint32_t phase, mcs;
    int sync_pass;
    uint64_t* claims;
    uint32_t sync_tag;
    int lblock;     // this block's index in its group
    int gtid, gsz;  // thread index and count in the group (fold)
};

// n queue slots for the calling wave (n the same on every active lane): the
// first active lane reserves them with one LDS atomic, and the base comes
// back to every lane through a scalar register (a cross-lane shuffle would
// cost a further LDS round trip).
__device__ __forceinline__ uint32_t wave_reserve(unsigned int* ctr, uint32_t n) {
    uint32_t base = 0;
    if (__lane_id() == (unsigned int)__builtin_amdgcn_readfirstlane((int)__lane_id())) base = atomicAdd(ctr, n);
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)base);
}

// Thread t of block b draws the attempts of same-colour sub-blocks
// (b ATT + r) blockDim + t for r < ATT (consecutive threads, consecutive
// sub-blocks), so the block's queue collects up to ATT blockDim attempts.
template <int D, bool STAGE, bool FULL, int ATT, class ST, bool XT = true>
__device__ __forceinline__ void compact_phase(const SweepConsts& k, const CompactPhase& a) {
    static_assert(!FULL || ATT == 1, "the full-neighbourhood queue holds one attempt per thread");
    __shared__ LdsParams sp;
    __shared__ QueuedAttempt<D, FULL, ST> queue[kBlock * ATT];
    __shared__ uint32_t nqueued;
    constexpr int NQ = QueuedAttempt<D, FULL, ST>::NQ;
    const int gtid = a.gtid;
    const int gsz = a.gsz;
    const PhaseTables T = tables_of(k, a.g);
    // Loads in flight together: parameters, the fold's deltas, the draws.
    constexpr int RP = (kParamWords + kBlock - 1) / kBlock;
    constexpr int RT = kThrLds / kBlock;
    const int npar = STAGE ? param_words(k.nt) : 0;
    const bool thr_lds = STAGE && k.nthr <= kThrLds;
    int32_t pv[RP];
    uint32_t tv[RT];
#pragma unroll
    for (int j = 0; j < RP; j++) {
        const int w = threadIdx.x + j * kBlock;
        pv[j] = w < npar ? k.prm->w[w] : 0;
    }
#pragma unroll
    for (int j = 0; j < RT; j++) {
        const int w = threadIdx.x + j * kBlock;
        tv[j] = thr_lds && w < k.nthr ? k.thr[w] : 0u;
    }
    const bool fold = a.sync_pass != 1;
    const int64_t dv = fold && gtid <= k.ncells ? T.Dp[gtid] : 0;
    using Lat = GlobalLat<D, false, ST>;
    const Lat lat{lat_base<ST>(k), k.lgL, (1 << k.lgL) - 1};
    // Draws (kernels2d.cu:199-243 in the shared roles).  When two sub-blocks
    // share a Philox block (attempt_words), rounds 2p and 2p + 1 of a thread
    // take the pair of sub-blocks 2t and 2t + 1 of their 2 blockDim range, so
    // the block is computed once.
    const bool paired = ATT % 2 == 0 && k.lc == 0 && a.lh > 0 && a.lstride == k.lstride;
    bool active[ATT];
    QueuedAttempt<D, FULL, ST> e[ATT];
    uint32_t wp[4];
#pragma unroll
    for (int r = 0; r < ATT; r++) {
        const int q = paired ? (a.lblock * ATT + (r & ~1)) * kBlock + 2 * (int)threadIdx.x + (r & 1)
                             : (a.lblock * ATT + r) * kBlock + (int)threadIdx.x;
        active[r] = false;
        if (q < a.nsub) {
            const int hm = (1 << a.lh) - 1;
            const int bx = ((q & hm) << a.lstride) + a.cx;
            const int by = (((q >> a.lh) & hm) << a.lstride) + a.cy;
            const int bz = D == 3 ? ((q >> (2 * a.lh)) << a.lstride) + a.cz : 0;
            const uint32_t sb = sub_block_id<D>(k, bx, by, bz);
            uint32_t w[4];
            if (!paired) {
                attempt_words(k, sb, bx, a.mcs, a.phase, 0u, w);
            } else if ((r & 1) == 0) {  // the even sub-block: its own counter
                philox4x32_10(sb, a.mcs, a.phase, 0u, k.k0, k.k1, wp);
                w[0] = wp[0];
                w[1] = wp[1];
                w[2] = wp[2];
            } else {  // q and q - 1 are both below nsub (nsub is even)
                w[0] = wp[0];
                w[1] = wp[3];
                w[2] = wp[0];
            }
            const int lc = k.lc;
            const uint32_t cm = (1u << lc) - 1u;
            constexpr int N = D == 3 ? 26 : 8;
            int tx = (bx << lc) + (int)(w[0] & cm);
            int ty = (by << lc) + (int)((w[0] >> lc) & cm);
            int tz = D == 3 ? (bz << lc) + (int)((w[0] >> (2 * lc)) & cm) : 0;
            int dir = (int)(((uint64_t)w[1] * (uint64_t)N) >> 32);
            dir += dir >= N / 2;
            if (k.roles) {  // non-shared roles: see draw_attempt
                int ox, oy, oz;
                block_offset<D>(dir, ox, oy, oz);
                const int Lm = (1 << k.lgL) - 1;
                tx = (tx + ox) & Lm;
                ty = (ty + oy) & Lm;
                tz = D == 3 ? (tz + oz) & Lm : 0;
                dir = N - dir;  // NB - 1 - dir
            }
            const typename Lat::Pos t = lat.at(tx, ty, tz);
            uint32_t vs, vt;
            if constexpr (NQ > 0) {
                lat.block(t, e[r].x);
                vs = 0;
#pragma unroll
                for (int n = 0; n < NQ; n++) vs = n == dir ? e[r].x[n] : vs;
                vt = e[r].x[NQ / 2];
            } else {
                int sx, sy, sz;
                block_offset<D>(dir, sx, sy, sz);
                vs = lat.nb(t, sx, sy, sz);
                vt = lat.self(t);
            }
            active[r] = vs != vt;  // kernels2d.cu:276
            e[r].site = t.i;
            e[r].vs = vs;
            e[r].w2 = w[2];
            e[r].aux = (uint32_t)dir;
            e[r].set_vt(vt);
        }
    }
#pragma unroll
    for (int j = 0; j < RP; j++) {
        const int w = threadIdx.x + j * kBlock;
        if (w < npar) sp.p[w] = pv[j];
    }
#pragma unroll
    for (int j = 0; j < RT; j++) {
        const int w = threadIdx.x + j * kBlock;
        if (thr_lds && w < k.nthr) sp.thr[w] = tv[j];
    }
    if (threadIdx.x == 0) nqueued = 0;
    __syncthreads();
    // Queue the active attempts: one LDS atomic per wave reserves the slots
    // of all its rounds.
    uint64_t masks[ATT];
    uint32_t total = 0;
#pragma unroll
    for (int r = 0; r < ATT; r++) {
        masks[r] = __ballot(active[r]);
        total += (uint32_t)__popcll(masks[r]);
    }
    if (total != 0) {
        uint32_t base = wave_reserve(&nqueued, total);
#pragma unroll
        for (int r = 0; r < ATT; r++) {
            const uint32_t before = __builtin_amdgcn_mbcnt_hi(
                (uint32_t)(masks[r] >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)masks[r], 0u));
            if (active[r]) queue[base + before] = e[r];
            base += (uint32_t)__popcll(masks[r]);
        }
    }
    __syncthreads();
    const int nq = (int)nqueued;
    for (int j = threadIdx.x; j < nq; j += kBlock) {
        const QueuedAttempt<D, FULL, ST>& q = queue[j];
        const int lgL = k.lgL, Lm = (1 << lgL) - 1;
        const typename Lat::Pos t{(int)(q.site & Lm), (int)((q.site >> lgL) & Lm),
                                  D == 3 ? (int)(q.site >> (2 * lgL)) : 0, q.site};
        Draw<D, Lat> d;
        d.w[0] = d.w[1] = d.w[3] = 0u;
        d.w[2] = q.w2;
        d.t = t;
        if constexpr (NQ > 0) {
#pragma unroll
            for (int n = 0; n < NQ; n++) d.nv[n] = q.x[n];
        } else {
            // The target holds the value the draw read (no other attempt of
            // the phase writes it): taking it from the queue keeps the
            // early exit on vs == vt from waiting for a reload.
            lat.block(t, d.nv);
            d.nv[(D == 3 ? 27 : 9) / 2] = q.vt();
        }
        d.vs = q.vs;
        d.dir = q.dir();
        d.tick = a.mcs;
        // The sub-block (cell_sync's key) from the drawn site: the target,
        // or with non-shared roles the source t + offset(dir).
        uint32_t sb = 0;
        if (a.sync_pass) {
            int ox = 0, oy = 0, oz = 0;
            if (k.roles) block_offset<D>(d.dir, ox, oy, oz);
            sb = sub_block_id<D>(k, ((t.x + ox) & Lm) >> k.lc, ((t.y + oy) & Lm) >> k.lc,
                                 D == 3 ? ((t.z + oz) & Lm) >> k.lc : 0);
        }
        const CellSync cy{a.sync_pass, a.claims, a.sync_tag};
        if constexpr (STAGE) {
            const LdsParamView P{sp.p, k.nt, k.nt * k.nt};
            finish_attempt<D, false, XT>(lat, k, P, thr_lds ? sp.thr : k.thr, T, cy, sb, d);
        } else {
            SmallParams P;
            P.load(k.prm);
            finish_attempt<D, false, XT>(lat, k, P, k.thr, T, cy, sb, d);
        }
    }
    if (fold) {
        if (dv != 0) {
            atomic_add_i64(T.W + gtid, dv);
            T.Dp[gtid] = 0;
        }
        for (int ci = gtid + gsz; ci <= k.ncells; ci += gsz) {
            const int64_t v = T.Dp[ci];
            if (v != 0) {
                atomic_add_i64(T.W + ci, v);
                T.Dp[ci] = 0;
            }
        }
    }
}


template <int D, bool STAGE, bool FULL, int ATT, class ST, bool XT>
__global__ __launch_bounds__(kBlock) void k_sweep_phase_compact(PhaseArgs a) {
    const CompactPhase cp{a.nsub, a.lh, a.lstride, a.cx, a.cy, a.cz, a.g, a.phase, a.mcs,
                          a.sync_pass, a.claims, a.sync_tag, (int)blockIdx.x,
                          (int)(blockIdx.x * blockDim.x + threadIdx.x),
                          (int)(gridDim.x * blockDim.x)};
    compact_phase<D, STAGE, FULL, ATT, ST, XT>(a.k, cp);
}

// Batched replicas, one colour phase, compacted: block b serves replica
// rep_of[b] as block b - block0 of its group.
template <int D, bool STAGE, int ATT, class ST>
__global__ __launch_bounds__(kBlock) void k_batch_phase_compact(BatchPhaseArgs a) {
    const int r = a.rep_of[blockIdx.x];
    const ReplicaDesc rd = a.reps[r];  // uniform per block: scalar loads
    const int lb = (int)blockIdx.x - rd.block0;
    const uint32_t m = a.m + (a.mbase ? *a.mbase : 0u);
    const CompactPhase cp{1 << (D * rd.lh), rd.lh, a.lstride, a.cx, a.cy, a.cz, a.g, a.phase,
                          rd.mcs0 + m, a.sync_pass, rd.claims, a.sync_tag, lb,
                          lb * (int)blockDim.x + (int)threadIdx.x, rd.nblocks * (int)blockDim.x};
    compact_phase<D, STAGE, false, ATT, ST>(rd.k, cp);
}

// Any number of types: parameters and thresholds staged in LDS.
template <int D, class ST>
__global__ __launch_bounds__(kBlock) void k_sweep_phase_lds(PhaseArgs a) {
    __shared__ LdsParams sp;
    LdsParamView P;
    const uint32_t* thr = stage_params(sp, a.k, &P);
    sweep_phase_body<D, ST>(a, P, thr);
}

// The reference stream: one thread per sub-block, staged parameters.
template <int D, class ST>
__global__ __launch_bounds__(kBlock) void k_sweep_phase_ref(PhaseArgs a) {
    __shared__ LdsParams sp;
    LdsParamView P;
    const uint32_t* thr = stage_params(sp, a.k, &P);
    sweep_phase_body<D, ST, true>(a, P, thr);
}

__global__ void k_ref_seed(int32_t* st, uint32_t n, int lr, int lR, int lgL, int d, uint32_t seed32) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t rm = (1u << lr) - 1u;
    const uint32_t x = (i & rm) << lR, y = ((i >> lr) & rm) << lR;
    const uint32_t z = d == 3 ? (i >> (2 * lr)) << lR : 0u;
    const uint32_t base = x | (y << lgL) | (z << (2 * lgL));  // kernels2d.cu:99
    st[i] = minstd_next((int32_t)(base + seed32));
}

hipError_t launch_ref_seed(int d, const SweepConsts& k, uint32_t seed32, hipStream_t s) {
    const uint32_t n = 1u << (d * k.lr);
    // a reference thread's first site: its index times positions_per_thread
    hipLaunchKernelGGL(k_ref_seed, dim3((n + 255) / 256), dim3(256), 0, s, k.rstate, n, k.lr,
                       k.lk + k.lc + k.lg, k.lgL, d, seed32);
    return hipGetLastError();
}

// Batched replicas, one colour phase: block b works for replica rep_of[b]
// on its same-colour sub-block (b - block0) * blockDim + t.
// REF: every replica draws from its reference stream (ref_attempts).
template <int D, class ST, bool REF = false>
__global__ __launch_bounds__(kBlock) void k_batch_phase(BatchPhaseArgs a) {
    const int r = a.rep_of[blockIdx.x];
    const ReplicaDesc rd = a.reps[r];  // uniform per block: scalar loads
    const SweepConsts& k = rd.k;
    __shared__ LdsParams sp;
    LdsParamView P;
    const uint32_t* thr = stage_params(sp, k, &P);
    const int gtid = (blockIdx.x - rd.block0) * blockDim.x + threadIdx.x;
    const int gsz = rd.nblocks * blockDim.x;
    const PhaseTables T = tables_of(k, a.g);
    const int lh = rd.lh;
    if (gtid < (1 << (D * lh))) {
        if (a.sync_pass == 3 || a.sync_pass == 5) rd.decide[gtid] = 0;
        const int hm = (1 << lh) - 1;
        int jx = gtid & hm, jy = (gtid >> lh) & hm, jz = D == 3 ? gtid >> (2 * lh) : 0;
        if (REF && k.lg) {  // (see ref_region)
            jx = ref_region(k, jx, a.gsx);
            jy = ref_region(k, jy, a.gsy);
            if (D == 3) jz = ref_region(k, jz, a.gsz);
        }
        const int bx = (jx << a.lstride) + a.cx;
        const int by = (jy << a.lstride) + a.cy;
        const int bz = D == 3 ? (jz << a.lstride) + a.cz : 0;
        const GlobalLat<D, false, ST> lat{lat_base<ST>(k), k.lgL, (1 << k.lgL) - 1};
        const CellSync cy{a.sync_pass, rd.claims, a.sync_tag, rd.decide, gtid};
        const uint32_t m = a.m + (a.mbase ? *a.mbase : 0u);
        if constexpr (REF)
            ref_attempts<D, false>(lat, k, P, thr, bx, by, bz, rd.mcs0 + m, T, cy, a.natt);
        else
            sub_block_attempts<D, false>(lat, k, P, thr, bx, by, bz, rd.mcs0 + m, a.phase, T, cy,
                                         a.att0, a.natt);
    }
    if (a.sync_pass != 1 && a.sync_pass != 3 && a.sync_pass != 5)
        fold_tables<false>(T, k.ncells, gtid, gsz);
}

// One 3D colour phase with positions_per_checkerboard = 1 (every sub-block
// is one site): block b stages the box of kBX x kBY x kBZ sites plus halo in
// LDS with coalesced loads and makes the attempts of the box's active sites
// (one per thread) against LDS.
constexpr int kBX = 16, kBY = 16, kBZ = 8;

__global__ __launch_bounds__(kBlock) void k_sweep_phase3_box(PhaseArgs a) {
    constexpr int SX = kBX + 2, SY = kBY + 2, SZ = kBZ + 2, SXY = SX * SY, NS = SXY * SZ;
    __shared__ LdsParams sp;
    __shared__ uint32_t box[NS];
    const SweepConsts& k = a.k;
    const int lgL = k.lgL, Lm = (1 << lgL) - 1;
    const int nbx = (Lm + 1) / kBX, nby = (Lm + 1) / kBY;
    const int b = blockIdx.x;
    const int x0 = (b % nbx) * kBX, y0 = ((b / nbx) % nby) * kBY, z0 = (b / (nbx * nby)) * kBZ;
    const uint32_t* __restrict__ gl = k.lat;
    // Box + halo: all loads issued before the LDS stores.
    constexpr int R = (NS + kBlock - 1) / kBlock;
    uint32_t v[R];
#pragma unroll
    for (int j = 0; j < R; j++) {
        const int i = threadIdx.x + j * kBlock;
        if (i < NS) {
            const int lz = i / SXY, rem = i - lz * SXY, ly = rem / SX, lx = rem - ly * SX;
            v[j] = gl[(uint32_t)((x0 + lx - 1) & Lm) | ((uint32_t)((y0 + ly - 1) & Lm) << lgL) |
                      ((uint32_t)((z0 + lz - 1) & Lm) << (2 * lgL))];
        }
    }
#pragma unroll
    for (int j = 0; j < R; j++) {
        const int i = threadIdx.x + j * kBlock;
        if (i < NS) box[i] = v[j];
    }
    LdsParamView P;
    const uint32_t* thr = stage_params(sp, k, &P);  // ends with __syncthreads
    // kBX/2 * kBY/2 * kBZ/2 = 256 active sites of this colour per box
    const int t = threadIdx.x;
    const int qx = t % (kBX / 2), qy = (t / (kBX / 2)) % (kBY / 2), qz = t / ((kBX / 2) * (kBY / 2));
    const LdsLat3 lat{box, k.lat, SX, SXY, x0, y0, z0, lgL};
    const PhaseTables T = tables_of(k, a.g);
    sub_block_attempts<3, false>(lat, k, P, thr, x0 + 2 * qx + a.cx, y0 + 2 * qy + a.cy,
                                 z0 + 2 * qz + a.cz, a.mcs, a.phase, T);
    fold_tables<false>(T, k.ncells, b * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}
static_assert((kBX / 2) * (kBY / 2) * (kBZ / 2) == kBlock, "one active site per thread");

// Barrier over nblocks co-resident blocks of a cooperative launch (the
// whole grid, or the blocks of one replica in a batch).  bar[0] counts
// arrivals, bar[1] is the generation; both persist in device memory between
// launches (bar[0] returns to 0 at every release).  Every word that crosses
// blocks goes through device-coherent accesses (ld/st<true>, atomics), so
// the barrier only has to make each wave's outstanding memory operations
// complete before its block arrives (s_waitcnt), with no cache maintenance.
__device__ __forceinline__ void wait_memory() {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// A waiting block gives up after this many polls (seconds; a phase takes
// microseconds): a block that never becomes resident, the one way the
// barrier can stall, then ends the launch with the error word set instead of
// hanging it.
constexpr long kBarrierSpinCap = 1L << 22;

// Returns false (the whole block) when the wait was abandoned; *err then
// holds 1 and every other waiter leaves at its next poll.
__device__ __forceinline__ bool grid_barrier(unsigned int* bar, unsigned int& gen,
                                             unsigned int nblocks, unsigned int* err) {
    __shared__ int abandoned;
    wait_memory();
    __syncthreads();
    if (threadIdx.x == 0) {
        int fail = 0;
        const unsigned int g = gen;
        const unsigned int prev =
            __hip_atomic_fetch_add(&bar[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (prev == nblocks - 1) {
            __hip_atomic_store(&bar[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            wait_memory();
            __hip_atomic_store(&bar[1], g + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            long polls = 0;
            while (__hip_atomic_load(&bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == g) {
                if (++polls > kBarrierSpinCap ||
                    __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) {
                    __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    fail = 1;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        wait_memory();
        abandoned = fail;
    }
    gen++;
    __syncthreads();
    return !abandoned;
}

// End of a persistent run: fold the last phase's deltas into the lagging
// table buffer (after a barrier), leaving both buffers equal.
__device__ __forceinline__ void settle_in_kernel(const SweepConsts& k, int G, int gtid, int gsz) {
    // After G phases tab[G & 1] is current and tab[(G + 1) & 1] lags by the
    // last phase's deltas del[(G + 1) & 1].
    int64_t* T = k.tab[(G + 1) & 1];
    int64_t* Dl = k.del[(G + 1) & 1];
    for (int ci = gtid; ci <= k.ncells; ci += gsz) {
        st<true>(T + ci, ld<true>(T + ci) + ld<true>(Dl + ci));
        st<true>(Dl + ci, (int64_t)0);
    }
}

// Whole run, lattice in global memory: the nblocks blocks of a group
// (lb = this block's index in it) stride over the same-colour sub-blocks of
// each phase and meet a group barrier between phases.
template <int D>
__device__ __forceinline__ void persist_global_run(const SweepConsts& k, uint32_t mcs0, int nmcs,
                                                   int lh, unsigned int* bar, unsigned int* err,
                                                   int lb, int nblocks, LdsParams& sp) {
    const int gtid = lb * blockDim.x + threadIdx.x;
    const int gsz = nblocks * blockDim.x;
    unsigned int gen = __hip_atomic_load(&bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    LdsParamView P;
    const uint32_t* thr = stage_params(sp, k, &P);
    const GlobalLat<D, true> lat{k.lat, k.lgL, (1 << k.lgL) - 1};
    const int hm = (1 << lh) - 1;
    const int nsub = 1 << (D * lh);
    constexpr int NCOL = 1 << D;
    int g = 0;
    for (int m = 0; m < nmcs; m++) {
        const uint32_t mcs = mcs0 + (uint32_t)m;
        for (int r = 0; r < k.rounds; r++) {
            for (int col = 0; col < NCOL; col++, g++) {
                if (g && !grid_barrier(bar, gen, nblocks, err)) return;
                int cx, cy, cz;
                colour_of<D>(col, cx, cy, cz);
                const PhaseTables T = tables_of(k, g);
                fold_tables<true>(T, k.ncells, gtid, gsz);
                const uint32_t phase = (uint32_t)(r * NCOL + col);
                for (int q = gtid; q < nsub; q += gsz) {
                    const int bx = 2 * (q & hm) + cx;
                    const int by = 2 * ((q >> lh) & hm) + cy;
                    const int bz = D == 3 ? 2 * (q >> (2 * lh)) + cz : 0;
                    sub_block_attempts<D, true>(lat, k, P, thr, bx, by, bz, mcs, phase, T);
                }
            }
        }
    }
    if (g && grid_barrier(bar, gen, nblocks, err)) settle_in_kernel(k, g, gtid, gsz);
}

template <int D>
__global__ __launch_bounds__(kBlock) void k_persist_global(PersistArgs a) {
    __shared__ LdsParams sp;
    persist_global_run<D>(a.k, a.mcs0, a.nmcs, a.lh, a.bar, a.bar + 2, blockIdx.x, gridDim.x, sp);
}

// Batched replicas in one cooperative launch: block b works for replica
// rep_of[b]; the blocks of one replica form a group with its own barrier,
// so replicas of different sizes advance independently.
template <int D>
__global__ __launch_bounds__(kBlock) void k_batch_global(BatchArgs a) {
    __shared__ LdsParams sp;
    const int r = a.rep_of[blockIdx.x];
    const ReplicaDesc rd = a.reps[r];  // uniform per block: scalar loads
    persist_global_run<D>(rd.k, rd.mcs0, a.nmcs, rd.lh, a.bars + 2 * r, a.err, blockIdx.x - rd.block0,
                          rd.nblocks, sp);
}

// Whole 2D run with each block's tile resident in LDS.  Block b owns the
// tile at ((b mod ntx) << ltx, (b / ntx) << lty); per phase it refreshes its
// halo ring from global memory (neighbouring tiles write through), makes the
// attempts of the tile's same-colour sub-blocks against LDS, and meets the
// grid barrier.
__global__ __launch_bounds__(kBlock) void k_persist_lds2d(PersistArgs a) {
    extern __shared__ uint32_t s[];
    const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const int gsz = gridDim.x * blockDim.x;
    const SweepConsts& k = a.k;
    unsigned int gen = __hip_atomic_load(&a.bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int TX = 1 << a.ltx, TY = 1 << a.lty, S = TX + 2;
    const int x0 = (int)(blockIdx.x & ((1u << a.lntx) - 1)) << a.ltx;
    const int y0 = (int)(blockIdx.x >> a.lntx) << a.lty;
    const int Lm = (1 << k.lgL) - 1;
    const uint32_t* __restrict__ gl = k.lat;
    __shared__ LdsParams sp;
    LdsParamView P;
    const uint32_t* thr = stage_params(sp, k, &P);
    // tile + halo
    for (int i = threadIdx.x; i < S * (TY + 2); i += blockDim.x) {
        const int ly = i / S, lx = i - ly * S;
        s[i] = gl[(uint32_t)((x0 + lx - 1) & Lm) | ((uint32_t)((y0 + ly - 1) & Lm) << k.lgL)];
    }
    __syncthreads();
    const LdsLat2 lat{s, k.lat, S, x0, y0, k.lgL};
    const int lq = a.ltx - k.lc - 1;  // log2 same-colour sub-blocks per tile row
    const int nq = 1 << (lq + a.lty - k.lc - 1);
    const int ring = 2 * S + 2 * TY;
    int g = 0;
    for (int m = 0; m < a.nmcs; m++) {
        const uint32_t mcs = a.mcs0 + (uint32_t)m;
        for (int r = 0; r < k.rounds; r++) {
            for (int col = 0; col < 4; col++, g++) {
                if (g) {
                    if (!grid_barrier(a.bar, gen, gridDim.x, a.bar + 2)) return;
                    for (int i = threadIdx.x; i < ring; i += blockDim.x) {
                        int lx, ly;
                        if (i < S) { lx = i; ly = 0; }
                        else if (i < 2 * S) { lx = i - S; ly = TY + 1; }
                        else if (i < 2 * S + TY) { lx = 0; ly = i - 2 * S + 1; }
                        else { lx = TX + 1; ly = i - 2 * S - TY + 1; }
                        s[ly * S + lx] = ld<true>(gl + ((uint32_t)((x0 + lx - 1) & Lm) |
                                                        ((uint32_t)((y0 + ly - 1) & Lm) << k.lgL)));
                    }
                    __syncthreads();
                }
                int cx, cy, cz;
                colour_of<2>(col, cx, cy, cz);
                const PhaseTables T = tables_of(k, g);
                fold_tables<true>(T, k.ncells, gtid, gsz);
                const uint32_t phase = (uint32_t)(r * 4 + col);
                for (int q = threadIdx.x; q < nq; q += blockDim.x) {
                    const int bx = (x0 >> k.lc) + 2 * (q & ((1 << lq) - 1)) + cx;
                    const int by = (y0 >> k.lc) + 2 * (q >> lq) + cy;
                    sub_block_attempts<2, true>(lat, k, P, thr, bx, by, 0, mcs, phase, T);
                }
            }
        }
    }
    if (g && grid_barrier(a.bar, gen, gridDim.x, a.bar + 2)) settle_in_kernel(k, g, gtid, gsz);
}

// A whole lattice resident in one workgroup's LDS (the single-workgroup
// form): sites in row-major order with periodic wrap.
template <int D, class T>
struct LdsWholeLat {
    using Site = T;
    T* s;
    int lgL, Lm;
    struct Pos {
        int x, y, z;
        uint32_t i;
    };
    __device__ __forceinline__ Pos at(int x, int y, int z) const {
        return {x, y, z, site_index<D>(x, y, z, lgL)};
    }
    __device__ __forceinline__ uint32_t self(const Pos& q) const { return s[q.i]; }
    __device__ __forceinline__ uint32_t nb(const Pos& q, int dx, int dy, int dz) const {
        return s[site_index<D>((q.x + dx) & Lm, (q.y + dy) & Lm, (q.z + dz) & Lm, lgL)];
    }
    __device__ __forceinline__ void put(const Pos& q, uint32_t v) const { s[q.i] = (T)v; }
    __device__ __forceinline__ uint32_t gindex(const Pos& q) const { return q.i; }
    __device__ __forceinline__ void block(const Pos& q, uint32_t* nv) const {
        constexpr int NB = D == 3 ? 27 : 9;
#pragma unroll
        for (int n = 0; n < NB; n++) {
            int ox, oy, oz;
            block_offset<D>(n, ox, oy, oz);
            nv[n] = nb(q, ox, oy, oz);
        }
    }
};

// Active sub-block offsets of phase pc of a round (as the host's
// phase_offsets): colour (cbX outermost, kernels2d.cu:179-180), then group.
template <int D>
__device__ __forceinline__ void phase_offsets_dev(const SweepConsts& k, int pc, int& ox, int& oy, int& oz) {
    const int K = k.K, S = k.S;
    const int ngrp = D == 3 ? S * S * S : S * S;
    const int col = pc / ngrp, grp = pc % ngrp;
    const int cx = D == 3 ? col / (K * K) : col / K;
    const int cy = D == 3 ? (col / K) % K : col % K;
    const int cz = D == 3 ? col % K : 0;
    const int gx = D == 3 ? grp / (S * S) : grp / S;
    const int gy = D == 3 ? (grp / S) % S : grp % S;
    const int gz = D == 3 ? grp % S : 0;
    ox = gx * K + cx;
    oy = gy * K + cy;
    oz = gz * K + cz;
}

// The single-workgroup form: one block of kSmallThreads threads runs a whole
// inner_iterations chunk of a lattice that fits in LDS (lattice, the
// snapshot table rows R and the phase deltas Dl), with __syncthreads in
// place of kernel boundaries.  Each phase: the attempts of the active
// sub-blocks (u = 1: drawn, the ones whose sites differ queued in LDS and
// finished densely; u > 1: one thread per sub-block, its u attempts in
// order), then the deltas are folded into R.  Same draws, roles, snapshot
// rule and integer arithmetic as the per-phase kernels, so the results are
// identical.  At the end the lattice and both table buffers are written back
// (the run's state is then settled).
constexpr int kSmallThreads = 1024;

template <int D, class ST, bool XT, class PV>
__device__ __forceinline__ void small_run(const SmallArgs& a, const PV& P, ST* lat_s, int64_t* R,
                                          int64_t* Dl, QueuedAttempt<D, false, ST>* queue,
                                          unsigned int* nqueued, const uint32_t* __restrict__ thr) {
    const SweepConsts& k = a.k;
    const int tid = threadIdx.x;
    const int nsub = 1 << (D * a.lh);
    const int hm = (1 << a.lh) - 1;
    const int Lm = (1 << k.lgL) - 1;
    using Lat = LdsWholeLat<D, ST>;
    const Lat lat{lat_s, k.lgL, Lm};
    const PhaseTables T{R, Dl, nullptr, nullptr};
    for (int m = 0; m < a.nmcs; m++) {
        const uint32_t mcs = a.mcs0 + (uint32_t)m;
        for (int r = 0; r < k.rounds; r++) {
            for (int pc = 0; pc < k.phases_per_round; pc++) {
                int ox, oy, oz;
                phase_offsets_dev<D>(k, pc, ox, oy, oz);
                const uint32_t phase = (uint32_t)(r * k.phases_per_round + pc);
                for (int q0 = 0; q0 < nsub; q0 += kSmallThreads) {
                    const int q = q0 + tid;
                    int bx = 0, by = 0, bz = 0;
                    if (q < nsub) {
                        bx = ((q & hm) << k.lstride) + ox;
                        by = (((q >> a.lh) & hm) << k.lstride) + oy;
                        bz = D == 3 ? ((q >> (2 * a.lh)) << k.lstride) + oz : 0;
                    }
                    if (k.u > 1) {
                        if (q < nsub) sub_block_attempts<D, false, XT>(lat, k, P, thr, bx, by, bz, mcs, phase, T);
                        continue;
                    }
                    // u = 1: draw, queue the attempts whose sites differ
                    bool active = false;
                    QueuedAttempt<D, false, ST> e;
                    if (q < nsub) {
                        Draw<D, Lat> d;
                        draw_attempt<D>(lat, k, sub_block_id<D>(k, bx, by, bz), bx, by, bz, mcs, phase, 0u, d);
                        const uint32_t vt = d.nv[(D == 3 ? 27 : 9) / 2];
                        active = d.vs != vt;  // kernels2d.cu:276
                        e.site = d.t.i;
                        e.vs = d.vs;
                        e.w2 = d.w[2];
                        e.aux = (uint32_t)d.dir;
                        e.set_vt(vt);
                    }
                    if (tid == 0) *nqueued = 0;
                    __syncthreads();
                    const uint64_t mask = __ballot(active);
                    if (mask) {
                        const uint32_t base = wave_reserve(nqueued, (uint32_t)__popcll(mask));
                        const uint32_t before = __builtin_amdgcn_mbcnt_hi(
                            (uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
                        if (active) queue[base + before] = e;
                    }
                    __syncthreads();
                    const int nq = (int)*nqueued;
                    for (int j = tid; j < nq; j += kSmallThreads) {
                        const QueuedAttempt<D, false, ST>& qa = queue[j];
                        const int lgL = k.lgL;
                        Draw<D, Lat> d;
                        d.t = typename Lat::Pos{(int)(qa.site & Lm), (int)((qa.site >> lgL) & Lm),
                                                D == 3 ? (int)(qa.site >> (2 * lgL)) : 0, qa.site};
                        d.w[0] = d.w[1] = d.w[3] = 0u;
                        d.w[2] = qa.w2;
                        lat.block(d.t, d.nv);
                        d.vs = qa.vs;
                        d.dir = qa.dir();
                        d.tick = mcs;
                        finish_attempt<D, false, XT>(lat, k, P, thr, T, CellSync{0, nullptr, 0}, 0u, d);
                    }
                    __syncthreads();  // (the queue is reused by the next round)
                }
                __syncthreads();
                // fold the phase's deltas into the snapshot
                for (int c = tid; c <= k.ncells; c += kSmallThreads) {
                    const int64_t v = Dl[c];
                    if (v) {
                        R[c] += v;
                        Dl[c] = 0;
                    }
                }
                __syncthreads();
            }
        }
    }
}

// XT = false: without the chemotaxis, act and persistence code (fewer
// registers, no scalar spills; the host picks it when those terms are off).
template <int D, class ST, bool XT>
__global__ __launch_bounds__(kSmallThreads) void k_small_lds(SmallArgs a) {
    extern __shared__ __align__(16) unsigned char small_smem[];
    const SweepConsts& k = a.k;
    const uint32_t sites = 1u << (D * k.lgL);
    const int rows = k.ncells + 1;
    int64_t* R = reinterpret_cast<int64_t*>(small_smem);
    int64_t* Dl = R + rows;
    auto* queue = reinterpret_cast<QueuedAttempt<D, false, ST>*>(Dl + rows);
    ST* lat_s = reinterpret_cast<ST*>(queue + kSmallThreads);
    // the thresholds after the lattice (an even number of sites: 4-byte aligned)
    uint32_t* thr_s = reinterpret_cast<uint32_t*>(lat_s + sites);
    __shared__ unsigned int nqueued;
    const ST* __restrict__ g = lat_base<ST>(k);
    for (uint32_t i = threadIdx.x; i < sites; i += kSmallThreads) lat_s[i] = g[i];
    if (a.thr_lds)
        for (int i = threadIdx.x; i < k.nthr; i += kSmallThreads) thr_s[i] = k.thr[i];
    const uint32_t* thr = a.thr_lds ? thr_s : k.thr;
    for (int c = threadIdx.x; c < rows; c += kSmallThreads) {
        R[c] = k.tab[0][c];  // settled: both buffers equal, deltas zero
        Dl[c] = 0;
    }
    __syncthreads();
    if (k.nt <= kSmallTypes) {
        SmallParams P;
        P.load(k.prm);
        small_run<D, ST, XT>(a, P, lat_s, R, Dl, queue, &nqueued, thr);
    } else {
        const LdsParamView P{k.prm->w, k.nt, k.nt * k.nt};  // (read from global memory)
        small_run<D, ST, XT>(a, P, lat_s, R, Dl, queue, &nqueued, thr);
    }
    ST* gw = lat_base<ST>(k);
    for (uint32_t i = threadIdx.x; i < sites; i += kSmallThreads) gw[i] = lat_s[i];
    for (int c = threadIdx.x; c < rows; c += kSmallThreads) {
        k.tab[0][c] = R[c];
        k.tab[1][c] = R[c];
    }
}

size_t small_lds_bytes(int d, int lgL, int ncells, bool narrow) {
    const size_t sites = (size_t)1 << (d * lgL);
    const size_t entry = narrow ? sizeof(QueuedAttempt<2, false, uint16_t>) : sizeof(QueuedAttempt<2, false, uint32_t>);
    const size_t entry3 = narrow ? sizeof(QueuedAttempt<3, false, uint16_t>) : sizeof(QueuedAttempt<3, false, uint32_t>);
    return 16 * ((size_t)ncells + 1) + (d == 3 ? entry3 : entry) * kSmallThreads + sites * (narrow ? 2 : 4);
}

hipError_t launch_small(int d, bool narrow, const SmallArgs& a, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
    const size_t bytes = small_lds_bytes(d, a.k.lgL, a.k.ncells, narrow) + (a.thr_lds ? 4 * (size_t)a.k.nthr : 0);
    const bool xt = a.k.chem || a.k.act || a.k.pers;
#define CPM_SMALL(DD, ST, XT)                                                                                        \
    do {                                                                                                             \
        hipError_t e = hipFuncSetAttribute((const void*)&k_small_lds<DD, ST, XT>,                                    \
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);                 \
        if (e != hipSuccess) return e;                                                                               \
        hipExtLaunchKernelGGL((k_small_lds<DD, ST, XT>), dim3(1), dim3(kSmallThreads), (unsigned)bytes, s, e0, e1,   \
                              0, a);                                                                                 \
    } while (0)
    if (d == 2) {
        if (narrow) {
            if (xt) CPM_SMALL(2, uint16_t, true); else CPM_SMALL(2, uint16_t, false);
        } else {
            if (xt) CPM_SMALL(2, uint32_t, true); else CPM_SMALL(2, uint32_t, false);
        }
    } else {
        if (narrow) {
            if (xt) CPM_SMALL(3, uint16_t, true); else CPM_SMALL(3, uint16_t, false);
        } else {
            if (xt) CPM_SMALL(3, uint32_t, true); else CPM_SMALL(3, uint32_t, false);
        }
    }
#undef CPM_SMALL
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// The resident form (2D, shared roles, c = 1, u = 1): one block per tile of
// TE x TE sites for a whole inner_iterations chunk.  The tile, a one-site
// halo and the table rows of every cell the tile may meet live in LDS;
// between colour phases a tile hears only from its eight neighbours, through
// mailbox words that carry their own phase tag.
//
// Holding cells.  Before the launch every cell gets the block of tiles its
// box, widened by one site (the halo) and by a margin for its growth within
// the launch (up to a quarter tile, less where a wide cell has no room),
// overlaps: one tile, two side by side or a 2 x 2 block (a cell wider than
// that fails the guard and the MCS run per phase).  Each tile holds the
// rows of its cells in LDS, slots in increasing cell order.  Every attempt
// whose cells a tile does not hold would read a row it cannot know: it sets
// the guard flag and the chunk is redone per phase from a copy.
//
// Why the rows stay exact.  A flip changes the rows of its two cells, and
// only tiles holding them make such flips.  The holders of a cell are
// neighbours of one another, so after every phase each tile sends each
// neighbour the deltas of the cells both hold (in increasing cell order,
// which both sides know) and adds the deltas it receives: every holder
// then has the row after the phase, which is what every attempt of the next
// phase must read.
//
// Why the halo stays exact.  Same-colour targets are two sites apart, so no
// attempt of a phase reads a site another attempt of it writes; the halo a
// tile needs in phase p is its neighbours' border after phase p - 1, which
// they send with the deltas.
//
// Mailboxes.  Tile b's packet to neighbour d after phase p sits in
// mail[b][p & 1][d]: its border sites facing d (two 16-bit sites or one
// 32-bit site per word), then one word per shared cell (perimeter delta
// << 12 plus area delta).  Each 64-bit word carries tag p + 1 in its high
// half, so a reader polls the words themselves: no fence, flag or second
// round trip.  A tile writes mail[b][p & 1] again only after its neighbours
// have sent phase p + 1, which they do after reading phase p.
//
// Order within a phase: the Philox draws and the attempts whose 3 x 3
// block lies inside the tile are made while the neighbours' packets are on
// their way; then the halo and the deltas arrive, the remaining attempts
// are drawn, all accepted flips are committed, the tile's packets are sent
// and its own deltas join the rows.
constexpr uint8_t kResMed = 0xFE;   // the medium's slot (row 0, never changed)
constexpr uint8_t kResNone = 0xFF;  // a cell the tile does not hold

template <int LTE>
struct ResGeom {
    static constexpr int TE = 1 << LTE;
    // LDS row stride: the tile and its halo need TE + 2; one more makes it
    // odd, so that same-colour rows (two apart) of 16-bit sites start on
    // banks of opposite parity
    static constexpr int S = TE + 3;
    static constexpr int NQ = TE * TE / 4;  // attempts per tile and phase
};

// Threads per tile of kernel variant VAR: one attempt each on the lean path
// (variants 0 and 3; at most 128 VGPRs: 16 waves per CU), two on the
// general one.
template <int LTE, int VAR>
constexpr int res_threads() {
    return (1 << (2 * LTE)) / (VAR == 0 || VAR == 3 ? 4 : 8);
}

// Neighbour direction d = 0..7: the 3 x 3 block of tiles without its centre,
// dx fastest; the opposite direction is 7 - d.
__device__ __forceinline__ void res_dir(int d, int& dx, int& dy) {
    const int k = d < 4 ? d : d + 1;
    dx = k % 3 - 1;
    dy = k / 3 - 1;
}

// Tile (tx, ty) holds the cell of holder word hw (res_holder_word).
__device__ __forceinline__ bool res_holds(uint32_t hw, int tx, int ty, int ntm) {
    if (!(hw >> 31)) return false;
    const int tx0 = (int)(hw & 0xFFFu), ty0 = (int)((hw >> 12) & 0xFFFu);
    return ((tx - tx0) & ntm) <= (int)((hw >> 24) & 1u) && ((ty - ty0) & ntm) <= (int)((hw >> 25) & 1u);
}

// The tile's lattice in LDS (sites and halo, row stride S).
template <class ST, int LTE>
struct ResLat {
    using Site = ST;
    ST* s;
    int x0, y0, lgL;
    struct Pos {
        int li;       // LDS index
        int lx, ly;   // 1 .. TE inside the tile
        uint32_t gi;  // global site index
    };
    __device__ __forceinline__ uint32_t self(const Pos& q) const { return s[q.li]; }
    __device__ __forceinline__ uint32_t nb(const Pos& q, int dx, int dy, int) const {
        return s[q.li + dy * ResGeom<LTE>::S + dx];
    }
    __device__ __forceinline__ void put(const Pos& q, uint32_t v) const { s[q.li] = (ST)v; }
    __device__ __forceinline__ void block(const Pos& q, uint32_t* nv) const { lds_block<2>(*this, q, nv); }
    __device__ __forceinline__ uint32_t gindex(const Pos& q) const { return q.gi; }
};

// The rows of the cells a tile holds: R the phase-start rows, Dl this
// phase's deltas, by slot.  Meeting a cell the tile does not hold sets *miss.
struct TileTables {
    const uint8_t* map;  // cell -> slot, kResNone: not held
    const int64_t* R;
    int64_t* Dl;
    unsigned int* miss;
    template <bool CO>
    __device__ __forceinline__ int64_t row(uint32_t c) const {
        const uint32_t s = map[c];
        if (s == kResNone) {
            *miss = 1u;
            return 0;
        }
        return R[s];
    }
    __device__ __forceinline__ void add(uint32_t c, int64_t delta) const {
        const uint32_t s = map[c];
        if (s != kResNone) atomicAdd(reinterpret_cast<unsigned long long*>(Dl + s), (unsigned long long)delta);
    }
};

// A queued attempt of the resident form (32 bytes of LDS).
constexpr uint32_t kResRing = 0x80000000u;  // li flag: the 3 x 3 block reaches the halo
struct ResEntry {
    uint32_t li;    // target (LDS index), | kResRing
    uint32_t vs, vt, w2;
    uint32_t meta;  // dir | on << 4 | as << 8 | at << 16
    uint32_t pad;
    int64_t e;      // the terms that need no rows
};

// The lean attempt path's parameters (at most four types) as LDS rows.
struct ResParams4 {
    int4 J[4];   // J[a][0..3]
    int4 JS[4];  // JS[a][0..3]
    int4 T[4];   // lambda_area, target area, lambda_perimeter, target perimeter of type a
};

// The queue in LDS, one array per field: consecutive lanes touch
// consecutive words (a 32-byte record per entry would put eight lanes on
// one bank).
template <int NQ>
struct ResQueue {
    uint32_t li[NQ], vs[NQ], vt[NQ], w2[NQ], meta[NQ];
    int64_t e[NQ];
    __device__ __forceinline__ ResEntry get(int i) const { return {li[i], vs[i], vt[i], w2[i], meta[i], 0u, e[i]}; }
};

template <class ST, int LTE, int NT>
struct ResShared {
    using G = ResGeom<LTE>;
    ST tile[G::S * G::S];
    // 1 where this phase's packets changed a halo site (the interior stays
    // 0), so a ring attempt prepared before they arrived is known to stand
    // (lean path, c = 1 with shared roles).
    uint8_t hdirty[G::S * G::S];
    int64_t R[256];
    int64_t D[2][256];           // deltas of the phases of each parity, by slot
    uint32_t cell[kResSlots];
    uint8_t nbm[kResSlots];      // bit d: neighbour d holds the cell too
    uint8_t dl[8][kResDirCap];   // slots shared with neighbour d, increasing cell order
    int dn[8];
    int woff[9];                 // first word of each packet in the block's word order
    uint32_t thr[kThrLds];
    int32_t prm[kParamWords];    // more than four types: the parameters, staged
    ResParams4 p4;               // the lean path's parameters
    int scan[NT];
    ResQueue<G::NQ> queue;       // this phase's attempts whose source and target differ, and the ring's
    uint4 draws[NT];             // the next phase's Philox blocks of each pair of same-colour sites, or staged candidates
    // Per phase parity: the queue length, and what stops the launch (1: an
    // attempt met a cell the tile does not hold; 2: a wait gave up), raised
    // in the phase before and read with the length after the first barrier.
    uint2 ctl[2];
    unsigned int quit;           // the set-up's guard
};

// Border site q (along the edge) facing (dx, dy), LDS coordinates 1 .. TE.
template <int LTE>
__device__ __forceinline__ void res_border(int dx, int dy, int q, int& lx, int& ly) {
    constexpr int TE = ResGeom<LTE>::TE;
    lx = dx > 0 ? TE : dx < 0 ? 1 : 1 + q;
    ly = dy > 0 ? TE : dy < 0 ? 1 : 1 + q;
}

// Mailbox words a thread reads and writes each phase: at most WPT of them.
template <class ST, int LTE, int NT>
struct ResWords {
    static constexpr int TE = ResGeom<LTE>::TE;
    static constexpr int SPW = 4 / (int)sizeof(ST);  // sites per word
    static constexpr int EDGE = TE / SPW;             // site words of an edge packet
    static constexpr int MAXW = 4 * (EDGE + kResDirCap) + 4 * (1 + kResDirCap);
    static constexpr int WPT = (MAXW + NT - 1) / NT;
    // A word's LDS side: sites (index of the first, stride to the next,
    // count) or a delta (slot).
    static __device__ __forceinline__ uint32_t sites(int li, int stride, int n) {
        return (uint32_t)li | ((uint32_t)stride << 16) | ((uint32_t)n << 28);
    }
    static __device__ __forceinline__ uint32_t delta(int slot) { return (uint32_t)slot | 0x80000000u; }
};

// One queued attempt of the resident form, in two parts: res_prepare()
// applies the fixed-type rules and evaluates the terms that need no cell
// rows (adhesion, chemotaxis); res_commit() adds the area and perimeter
// terms from the phase-start rows, accepts and commits.  The split lets a
// tile prepare its interior attempts while its neighbours' packets travel.
// On entry vs and vt hold the drawn source and target values, which differ.
template <class ST, int LTE, bool XT, class PV>
__device__ __forceinline__ void res_prepare(const ResLat<ST, LTE>& lat, const SweepConsts& k, const PV& P,
                                            int x0, int y0, ResEntry& q) {
    constexpr int S = ResGeom<LTE>::S;
    const uint32_t li = q.li & ~kResRing, vt = q.vt;
    const int dir = (int)(q.meta & 0xFu);
    uint32_t vs = q.vs;
    unsigned long long* const ctr = k.counters;
    if (ctr) count_attempts(ctr, 0);
    if (k.fixed_mask) {  // fixed types (kernels2d.cu:281-290); the act term is off in this form
        if ((k.fixed_mask >> (vt >> k.idbits)) & 1u) return;
        if ((k.fixed_mask >> (vs >> k.idbits)) & 1u) {
            vs = 0;
            if (vs == vt) return;
        }
    }
    if (ctr) count_attempts(ctr, 1);  // kernels2d.cu:292
    const int ly = (int)li / S, lx = (int)li - ly * S;
    const typename ResLat<ST, LTE>::Pos t{(int)li, lx, ly,
                                          (uint32_t)(x0 + lx - 1) | ((uint32_t)(y0 + ly - 1) << lat.lgL)};
    uint32_t nv[9];
    lat.block(t, nv);
    const int ts = (int)(vs >> k.idbits), tt = (int)(vt >> k.idbits);
    int as, at;
    int64_t e = adhesion_energy<2>(k, P, nv, vs, vt, ts, tt, as, at);
    if (XT && k.chem) e += chem_energy<2>(k, lat, t, dir, ts, tt);
    q.vs = vs;
    q.e = e;
    q.meta = (uint32_t)dir | (1u << 4) | ((uint32_t)as << 8) | ((uint32_t)at << 16);
}

template <class ST, int LTE, class PV>
__device__ __forceinline__ void res_commit(ST* tile, const SweepConsts& k, const PV& P, const TileTables& T,
                                           const uint32_t* thr_s, bool thr_lds, const ResEntry& q) {
    const uint32_t meta = q.meta;
    if (!((meta >> 4) & 1u)) return;
    const int as = (int)((meta >> 8) & 0xFFu), at = (int)(meta >> 16);
    const uint32_t vs = q.vs, vt = q.vt;
    const int ts = (int)(vs >> k.idbits), tt = (int)(vt >> k.idbits);
    const uint32_t cs = vs & k.idmask, ct = vt & k.idmask;
    const int64_t row_s = cs ? T.row<false>(cs) : 0;
    const int64_t row_t = ct ? T.row<false>(ct) : 0;
    const int64_t dh = q.e + row_energy<2>(P, cs, ct, ts, tt, row_s, row_t, as, at);
    bool accept = dh <= 0;
    if (!accept && dh < k.nthr) {
        // (a device-coherent load for the long table keeps the two loads
        // apart, so the LDS one stays an LDS instruction)
        const uint32_t th = thr_lds ? thr_s[dh] : ld<true>(k.thr + dh);
        accept = q.w2 < th;
    }
    if (!accept) return;
    unsigned long long* const ctr = k.counters;
    if (ctr) {
        count_attempts(ctr, 2);  // kernels2d.cu:526
        count_attempts(ctr, 3);  // kernels2d.cu:531
    }
    tile[q.li & ~kResRing] = (ST)vs;  // kernels2d.cu:548
    // area and perimeter deltas (kernels2d.cu:552-569)
    constexpr int N = 8;
    if (cs) T.add(cs, (int64_t)(N - 2 * as) * 4294967296LL + 1);
    if (ct) T.add(ct, (int64_t)(2 * at - N) * 4294967296LL - 1);
}

// The lean attempt path of the resident form: at most four types, no fixed
// types, no chemotaxis, no counters, a threshold table that fits LDS.  The
// parameters are LDS rows (one 16-byte load each), the neighbour counts are
// packed in one word and both cells' terms are computed without branches,
// which shortens the dependent chain each phase waits on.
__device__ __forceinline__ int32_t pick4(int i, const int4& v) { return pick4(i, v.x, v.y, v.z, v.w); }

// The lean path's meta word: dir | on << 4 | as << 8 | at << 12 | the
// source's slot << 16 | the target's slot << 24 (slots as the tile's map
// gives them, fixed for a launch, so the prepare looks them up).
template <class ST, int LTE>
__device__ __forceinline__ void res_prepare_lean(const ST* tile, const uint8_t* map, const ResParams4& pr,
                                                 uint32_t idbits, uint32_t idmask, ResEntry& q) {
    constexpr int S = ResGeom<LTE>::S;
    const int li = (int)(q.li & ~kResRing);
    const uint32_t vs = q.vs, vt = q.vt;
    const int ts = (int)(vs >> idbits), tt = (int)(vt >> idbits);
    const int4 jr = pr.J[ts], jsr = pr.JS[tt];
    int32_t j0 = jr.x, j1 = jr.y, j2 = jr.z, j3 = jr.w, k0 = jsr.x, k1 = jsr.y, k2 = jsr.z, k3 = jsr.w;
    // (the diagonal entries as loads of their own: a select among a row's
    // four loaded words compiles to branches)
    int32_t jss = reinterpret_cast<const int32_t*>(&pr.J[ts])[ts];
    int32_t jtt = reinterpret_cast<const int32_t*>(&pr.JS[tt])[tt];
    uint32_t ss = map[vs & idmask], st_ = map[vt & idmask];
    uint32_t nv[8] = {tile[li - S - 1], tile[li - S], tile[li - S + 1], tile[li - 1],
                      tile[li + 1],     tile[li + S - 1], tile[li + S], tile[li + S + 1]};
    // One wait for every load above: the statement takes and gives back
    // each value, so nothing computed from them starts before all have
    // arrived (the scheduler otherwise interleaves arithmetic with the
    // loads and waits three times).
    asm volatile(""
                 : "+v"(nv[0]), "+v"(nv[1]), "+v"(nv[2]), "+v"(nv[3]), "+v"(nv[4]), "+v"(nv[5]), "+v"(nv[6]),
                   "+v"(nv[7]), "+v"(ss), "+v"(st_), "+v"(jss), "+v"(jtt), "+v"(j0), "+v"(j1), "+v"(j2), "+v"(j3),
                   "+v"(k0), "+v"(k1), "+v"(k2), "+v"(k3));
    uint32_t hist = 0;
    int as = 0, at = 0;
#pragma unroll
    for (int n = 0; n < 8; n++) {
        hist += 1u << ((nv[n] >> idbits) * 5u);
        as += nv[n] == vs;
        at += nv[n] == vt;
    }
    // adhesion (kernels2d.cu:359-402,477) from the counts per type
    int32_t adh = (j0 - k0) * (int32_t)(hist & 31u) + (j1 - k1) * (int32_t)((hist >> 5) & 31u) +
                  (j2 - k2) * (int32_t)((hist >> 10) & 31u) + (j3 - k3) * (int32_t)((hist >> 15) & 31u);
    adh += at * jtt - as * jss;
    q.e = adh;
    q.meta = (q.meta & 0xFu) | (1u << 4) | ((uint32_t)as << 8) | ((uint32_t)at << 12) | (ss << 16) | (st_ << 24);
}

// Perimeter and area terms of one cell (kernels2d.cu:479-520): sign +1 for
// the gaining source cell, -1 for the losing target cell.
__device__ __forceinline__ int64_t res_cell_terms(const int4& tp, int64_t row, int64_t dP, int sgn) {
    const int64_t A = (int64_t)(uint32_t)row, Pm = row >> 32;
    return (int64_t)tp.x * (1 + sgn * (2 * A - 2 * (int64_t)tp.y)) +
           (int64_t)tp.z * dP * (dP + 2 * (Pm - tp.w));
}

template <class ST, int LTE>
__device__ __forceinline__ void res_commit_lean(ST* tile, const int64_t* R, int64_t* Dl, const uint32_t* thr,
                                                int nthr, const ResParams4& pr, uint32_t idbits, uint32_t idmask,
                                                const ResEntry& q, bool& miss) {
    const uint32_t meta = q.meta;
    if (!((meta >> 4) & 1u)) return;
    constexpr int N = 8;
    const int as = (int)((meta >> 8) & 0xFu), at = (int)((meta >> 12) & 0xFu);
    const uint32_t vs = q.vs, vt = q.vt;
    const int ts = (int)(vs >> idbits), tt = (int)(vt >> idbits);
    const uint32_t cs = vs & idmask, ct = vt & idmask;
    const uint32_t ss = (meta >> 16) & 0xFFu, st_ = meta >> 24;  // (the medium has slot kResMed, row 0)
    const int4 ps = pr.T[ts], pt = pr.T[tt];
    const int64_t rs = R[ss], rt = R[st_];
    const int64_t dPs = N - 2 * as, dPt = 2 * at - N;
    int64_t dh = q.e;
    dh += cs ? res_cell_terms(ps, rs, dPs, 1) : 0;
    dh += ct ? res_cell_terms(pt, rt, dPt, -1) : 0;
    const int64_t ix = dh < 0 ? 0 : (dh >= nthr ? nthr - 1 : dh);
    const uint32_t th = thr[ix];
    const bool accept = dh <= 0 || (dh < nthr && q.w2 < th);
    miss |= ss == kResNone || st_ == kResNone;
    if (!accept || ss == kResNone || st_ == kResNone) return;
    tile[q.li & ~kResRing] = (ST)vs;  // kernels2d.cu:548
    // area and perimeter deltas (kernels2d.cu:552-569)
    if (cs) atomicAdd(reinterpret_cast<unsigned long long*>(Dl + ss), (unsigned long long)(dPs * 4294967296LL + 1));
    if (ct) atomicAdd(reinterpret_cast<unsigned long long*>(Dl + st_), (unsigned long long)(dPt * 4294967296LL - 1));
}

// The Philox block of pair q (same-colour row q / (TE / 4), x-neighbouring
// pair q mod (TE / 4)) of a tile with origin (x0, y0) in phase p of a
// chunk starting at MCS mcs0: the counter of the pair's even site, as
// attempt_words draws it.
template <int LTE>
__device__ __forceinline__ uint4 res_pair_draws(const SweepConsts& k, int x0, int y0, uint32_t mcs0,
                                                unsigned int p, int q) {
    constexpr int TE = ResGeom<LTE>::TE;
    const int col = (int)(p & 3u);
    int cx, cy, cz;
    colour_of<2>(col, cx, cy, cz);
    const int j = q / (TE / 4), jc = q % (TE / 4);
    const int bx = x0 + 4 * jc + cx, by = y0 + 2 * j + cy;
    uint32_t wp[4];
    philox4x32_10((uint32_t)bx | ((uint32_t)by << k.lgL), mcs0 + (p >> 2), (uint32_t)col, 0u, k.k0, k.k1, wp);
    return make_uint4(wp[0], wp[1], wp[2], wp[3]);
}

// The schedule of phase p of a launch, any schedule with u = 1 (every count
// a power of two, so shifts and masks): its MCS and Philox phase index
// (round * phases_per_round + phase in the round, as phase_offsets_dev and
// the per-phase kernels count them), and for a tile at (x0, y0) the active
// sub-blocks whose targets may lie in it, origins fx + ix * sbs, fy + iy *
// sbs (sites) for ix < ncx, iy < ncy; candidate ci = iy * ncx + ix.
struct ResSched {
    uint32_t mcs, phase;
    int fx, fy, ncx, ncy, lsbs;
};

template <int LTE>
__device__ __forceinline__ ResSched res_sched(const SweepConsts& k, uint32_t mcs0, unsigned int p, int x0,
                                              int y0) {
    constexpr int TE = 1 << LTE;
    const int lK = __builtin_ctz((unsigned)k.K), lS = __builtin_ctz((unsigned)k.S);
    const int lppr = 2 * (lK + lS), lper = __builtin_ctz((unsigned)k.rounds) + lppr;
    ResSched r;
    r.mcs = mcs0 + (p >> lper);
    r.phase = p & ((1u << lper) - 1u);
    const uint32_t pc = r.phase & ((1u << lppr) - 1u);
    const uint32_t col = pc >> (2 * lS), grp = pc & ((1u << (2 * lS)) - 1u);
    const int ox = ((int)(grp >> lS) << lK) + (int)(col >> lK);
    const int oy = ((int)(grp & (uint32_t)(k.S - 1)) << lK) + (int)(col & (uint32_t)(k.K - 1));
    const int lc = k.lc, c = 1 << lc;
    r.lsbs = lc + k.lstride;
    const int sbm = (1 << r.lsbs) - 1;
    r.fx = x0 - c + ((ox * c - x0 + c) & sbm);
    r.fy = y0 - c + ((oy * c - y0 + c) & sbm);
    r.ncx = max(0, ((x0 + TE - r.fx) >> r.lsbs) + 1);
    r.ncy = max(0, ((y0 + TE - r.fy) >> r.lsbs) + 1);
    return r;
}

// Candidate ci of a phase (ci < ncx ncy) as the tile at (x0, y0) stages it:
// its sub-block's attempt (attempt_words: position, direction and roles as
// the per-phase kernels draw them) and, when the target lies in the tile at
// (lx, ly) (1 .. TE), (1 << 31 | dir << 16 | ly << 8 | lx, w2); (0, 0)
// otherwise.
template <int LTE>
__device__ __forceinline__ uint2 res_candidate_target(const SweepConsts& k, const ResSched& r, int ci, int Lm,
                                                      int x0, int y0) {
    constexpr int TE = 1 << LTE;
    const int iy = ci / r.ncx, ix = ci - iy * r.ncx;
    const int bx = ((r.fx + (ix << r.lsbs)) & Lm) >> k.lc;
    const int by = ((r.fy + (iy << r.lsbs)) & Lm) >> k.lc;
    uint32_t w[4];
    attempt_words(k, sub_block_id<2>(k, bx, by, 0), bx, r.mcs, r.phase, 0u, w);
    const int lc = k.lc, cm = (1 << lc) - 1;
    int px = (bx << lc) + (int)(w[0] & (uint32_t)cm);
    int py = (by << lc) + (int)((w[0] >> lc) & (uint32_t)cm);
    int dir = (int)(((uint64_t)w[1] * 8ull) >> 32);
    dir += dir >= 4;
    if (k.roles) {  // the drawn site is the source, its neighbour the target
        int ddx, ddy, ddz;
        block_offset<2>(dir, ddx, ddy, ddz);
        px = (px + ddx) & Lm;
        py = (py + ddy) & Lm;
        dir = 8 - dir;
    }
    const int lx = ((px - x0) & Lm) + 1, ly = ((py - y0) & Lm) + 1;
    return lx <= TE && ly <= TE
               ? make_uint2((1u << 31) | ((uint32_t)dir << 16) | ((uint32_t)ly << 8) | (uint32_t)lx, w[2])
               : make_uint2(0u, 0u);
}

// GEN: 0 for c = 1 with shared roles (pairs of same-colour sites), 1 for any
// other schedule (candidate sub-blocks), -1 to take it from the arguments.
template <class ST, int LTE, int NT, bool LEAN, bool XT, int GEN, class PV>
__device__ __forceinline__ void resident_run(const ResidentArgs& a, const PV& P, ResShared<ST, LTE, NT>& sh,
                                             uint8_t* map) {
    using G = ResGeom<LTE>;
    using WD = ResWords<ST, LTE, NT>;
    constexpr int TE = G::TE, S = G::S, SPW = WD::SPW, WPT = WD::WPT;
    constexpr int APT = G::NQ / NT;          // attempts per thread and phase (c = 1, shared roles)
    constexpr int NPAIR = G::NQ / 2;         // pairs of same-colour sites sharing a Philox block
    const SweepConsts& k = a.k;
    const bool gen = GEN < 0 ? a.gen != 0 : GEN == 1;
    // The ring's attempts prepared before the packets arrive (lean path, c = 1
    // with shared roles; measured slower for the other schedules)
    constexpr bool EARLY = LEAN && GEN == 0;
    const int tid = (int)threadIdx.x;
    const int b = (int)blockIdx.x;
    const int ntm = (1 << a.lnt) - 1;
    const int tx = b & ntm, ty = b >> a.lnt;
    const int x0 = tx << LTE, y0 = ty << LTE;
    const int Lm = (1 << k.lgL) - 1;
    ST* glat = lat_base<ST>(k);
    if (*a.bad != 0u) return;  // a cell wider than two tiles (set before the launch)
    for (int i = tid; i < S * S; i += NT) {
        const int ly = i / S, lx = i - ly * S;
        sh.tile[i] = glat[(uint32_t)((x0 + lx - 1) & Lm) | ((uint32_t)((y0 + ly - 1) & Lm) << k.lgL)];
        sh.hdirty[i] = 0;
    }
    const bool thr_lds = k.nthr <= kThrLds;
    if (thr_lds)
        for (int i = tid; i < k.nthr; i += NT) sh.thr[i] = k.thr[i];
    for (int c = tid; c <= k.ncells; c += NT) map[c] = c ? kResNone : kResMed;
    if (tid < 2) sh.D[tid][kResMed] = 0;
    if (tid == 0) sh.R[kResMed] = 0;
    if (LEAN && tid < 48) {
        const int f = tid / 16, a = (tid / 4) & 3, c = tid & 3;  // J, JS rows; then the terms by type
        int32_t* dst = reinterpret_cast<int32_t*>(&sh.p4);
        const int32_t* s4 = k.prm->s4;
        dst[tid] = f < 2 ? s4[tid] : s4[2 * kSmallTypes * kSmallTypes + c * kSmallTypes + a];
    }
    // slots in increasing cell order: thread t counts the cells of its range
    // it holds, a block scan gives each range its first slot
    const int per = (k.ncells + NT - 1) / NT;
    const int c0 = 1 + tid * per, c1 = min(c0 + per, k.ncells + 1);
    int cnt = 0;
    for (int c = c0; c < c1; c++) cnt += res_holds(a.holders[c], tx, ty, ntm);
    sh.scan[tid] = cnt;
    if (tid == 0) sh.quit = 0;
    __syncthreads();
    for (int off = 1; off < NT; off <<= 1) {
        const int v = tid >= off ? sh.scan[tid - off] : 0;
        __syncthreads();
        sh.scan[tid] += v;
        __syncthreads();
    }
    const int nslots = sh.scan[NT - 1];
    if (nslots > kResSlots) {  // more cells than slots: the guard
        if (tid == 0) atomicOr(a.bad, kResGuardSlots);
        return;
    }
    int slot = sh.scan[tid] - cnt;
    for (int c = c0; c < c1; c++) {
        const uint32_t hw = a.holders[c];
        if (!res_holds(hw, tx, ty, ntm)) continue;
        uint32_t m = 0;
#pragma unroll
        for (int d = 0; d < 8; d++) {
            int dx, dy;
            res_dir(d, dx, dy);
            m |= (uint32_t)res_holds(hw, (tx + dx) & ntm, (ty + dy) & ntm, ntm) << d;
        }
        map[c] = (uint8_t)slot;
        sh.cell[slot] = (uint32_t)c;
        sh.nbm[slot] = (uint8_t)m;
        sh.R[slot] = k.tab[0][c];
        sh.D[0][slot] = 0;
        sh.D[1][slot] = 0;
        slot++;
    }
    __syncthreads();
    if (tid < 8) {
        int n = 0;
        for (int s = 0; s < nslots; s++)
            if ((sh.nbm[s] >> tid) & 1u) {
                if (n < kResDirCap) sh.dl[tid][n] = (uint8_t)s;
                n++;
            }
        sh.dn[tid] = n;
    }
    __syncthreads();
    if (tid == 0) {
        int o = 0;
        unsigned int over = 0;
        for (int d = 0; d < 8; d++) {
            int dx, dy;
            res_dir(d, dx, dy);
            sh.woff[d] = o;
            over |= sh.dn[d] > kResDirCap;
            o += (dx != 0 && dy != 0 ? 1 : WD::EDGE) + sh.dn[d];
        }
        sh.woff[8] = o;
        sh.quit = over;
    }
    __syncthreads();
    if (sh.quit) {  // more shared cells than a packet holds: the guard
        if (tid == 0) atomicOr(a.bad, kResGuardShared);
        return;
    }
    // This thread's mailbox words: word offsets in the neighbour's packet
    // (read) and in this tile's own (written), without the parity, and the
    // LDS side of each.
    const uint32_t mstride = (uint32_t)res_mail_stride(TE);
    uint32_t wsrc[WPT], wdst[WPT], win[WPT], wout[WPT];
    uint32_t need = 0;
#pragma unroll
    for (int j = 0; j < WPT; j++) {
        const int w = tid + j * NT;
        wsrc[j] = wdst[j] = win[j] = wout[j] = 0;
        if (w < sh.woff[8]) {
            int d = 0;
            while (d < 7 && w >= sh.woff[d + 1]) d++;
            int dx, dy;
            res_dir(d, dx, dy);
            const int i = w - sh.woff[d];
            const int nb = ((tx + dx) & ntm) | (((ty + dy) & ntm) << a.lnt);
            wsrc[j] = ((uint32_t)nb * 16u + (uint32_t)(7 - d)) * mstride + (uint32_t)i;
            wdst[j] = ((uint32_t)b * 16u + (uint32_t)d) * mstride + (uint32_t)i;
            const bool corner = dx != 0 && dy != 0;
            if (i < (corner ? 1 : WD::EDGE)) {
                const int n = corner ? 1 : SPW, stride = dx == 0 ? 1 : S;
                int lx, ly;
                res_border<LTE>(-dx, -dy, i * SPW, lx, ly);  // the neighbour's border: this tile's halo
                win[j] = WD::sites((ly + dy * TE) * S + lx + dx * TE, stride, n);
                res_border<LTE>(dx, dy, i * SPW, lx, ly);  // this tile's border facing d
                wout[j] = WD::sites(ly * S + lx, stride, n);
            } else {
                win[j] = wout[j] = WD::delta(sh.dl[d][i - (corner ? 1 : WD::EDGE)]);
            }
            need |= 1u << j;
        }
    }
    // the halo sites of each word as last received (loaded at the start)
    uint32_t prev[WPT];
#pragma unroll
    for (int jj = 0; jj < WPT; jj++) {
        prev[jj] = 0;
        const uint32_t x = win[jj];
        if (EARLY && (need >> jj & 1u) && !(x >> 31)) {
            const int li = (int)(x & 0xFFFFu), stride = (int)((x >> 16) & 0xFFFu);
            prev[jj] = (uint32_t)sh.tile[li];
            if (SPW == 2 && (x >> 28) > 1u) prev[jj] |= (uint32_t)sh.tile[li + stride] << 16;
        }
    }
    const ResLat<ST, LTE> lat{sh.tile, x0, y0, k.lgL};
    if (tid == 0) sh.ctl[0] = sh.ctl[1] = make_uint2(0u, 0u);
    __syncthreads();
    // this thread's pair: same-colour row j, x-neighbouring pair jc
    // this thread's pair of same-colour sites (c = 1, shared roles): row j of
    // the colour, x-neighbouring pair jc; with one attempt per thread, the
    // pair's site h
    const int pair = APT == 2 ? tid : tid >> 1, hs = APT == 2 ? 0 : tid & 1;
    const int j = pair / (TE / 4), jc = pair % (TE / 4);
    const unsigned int Gp = (unsigned int)(a.nmcs * k.rounds * k.phases_per_round);
    if (!gen) {
        for (int q = tid; q < NPAIR; q += NT) sh.draws[q] = res_pair_draws<LTE>(k, x0, y0, a.mcs0, 0u, q);
    } else {
        const ResSched r = res_sched<LTE>(k, a.mcs0, 0u, x0, y0);
        if (tid < r.ncx * r.ncy) {
            const uint2 t = res_candidate_target<LTE>(k, r, tid, Lm, x0, y0);
            sh.draws[tid] = make_uint4(t.x, t.y, 0u, 0u);
        }
    }
    // (with one attempt per thread, thread t draws with pair t / 2, which
    // another wave wrote)
    __syncthreads();
    for (unsigned int p = 0;; p++) {
        const bool work = p < Gp;  // p == Gp: receive the last phase's deltas only
        // (the neighbours' packets of phase p - 1, halo sites and deltas, are
        // read from this parity)
        const uint32_t par_in = ((p - 1) & 1u) * 8u * mstride;
        unsigned long long v[WPT];
        ResEntry own{};  // this thread's first attempt of the phase (queue entry tid)
        uint32_t got = p > 0 ? 0u : need;
        unsigned int* const nq = &sh.ctl[p & 1u].x;
        unsigned int* const stop = &sh.ctl[(p + 1) & 1u].y;  // (raised in this phase)
        // Queue one or two attempts per lane: one LDS atomic per wave.
        auto enqueue = [&](const ResEntry* e, const bool* act) {
            const uint64_t m0 = __ballot(act[0]), m1 = __ballot(act[1]);
            const uint32_t total = (uint32_t)(__popcll(m0) + __popcll(m1));
            if (total != 0) {
                const uint32_t base = wave_reserve(nq, total);
                const uint32_t b0 = __builtin_amdgcn_mbcnt_hi((uint32_t)(m0 >> 32),
                                                              __builtin_amdgcn_mbcnt_lo((uint32_t)m0, 0u));
                const uint32_t b1 = __builtin_amdgcn_mbcnt_hi((uint32_t)(m1 >> 32),
                                                              __builtin_amdgcn_mbcnt_lo((uint32_t)m1, 0u));
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    if (!act[h]) continue;
                    const uint32_t i = base + (h ? (uint32_t)__popcll(m0) + b1 : b0);
                    sh.queue.li[i] = e[h].li;
                    sh.queue.vs[i] = e[h].vs;
                    sh.queue.vt[i] = e[h].vt;
                    sh.queue.w2[i] = e[h].w2;
                    sh.queue.meta[i] = e[h].meta;
                }
            }
        };
        // An attempt on target (lx, ly) of the tile with source direction
        // dir: the ring's (its 3 x 3 block reaches the halo) are queued as
        // they are, the others when their sites differ.
        auto stage = [&](int lx, int ly, int dir, uint32_t w2, ResEntry& e, bool& act) {
            const bool ring = lx == 1 || lx == TE || ly == 1 || ly == TE;
            const int li = ly * S + lx;
            e.li = (uint32_t)li | (ring ? kResRing : 0u);
            e.w2 = w2;
            e.meta = (uint32_t)dir;
            if (ring) {
                act = true;
            } else {
                int ox, oy, oz;
                block_offset<2>(dir, ox, oy, oz);
                e.vt = sh.tile[li];
                e.vs = sh.tile[li + oy * S + ox];
                act = e.vs != e.vt;  // kernels2d.cu:276
            }
        };
        if (work && !gen) {
            // c = 1, shared roles: each thread takes an x-neighbouring pair of
            // same-colour sites, whose Philox block was made in the previous
            // phase
            const int col = (int)(p & 3u);
            int cx, cy, cz;
            colour_of<2>(col, cx, cy, cz);
            const uint4 dw = sh.draws[pair];
            const uint32_t wp[4] = {dw.x, dw.y, dw.z, dw.w};
            const int ly = 2 * j + cy + 1;
            ResEntry e[2];
            bool act[2] = {false, false};
#pragma unroll
            for (int i = 0; i < APT; i++) {
                const int h = APT == 2 ? i : hs;
                int dir = (int)(((uint64_t)(h ? wp[3] : wp[1]) * 8ull) >> 32);
                dir += dir >= 4;
                stage(4 * jc + cx + 2 * h + 1, ly, dir, h ? wp[0] : wp[2], e[i], act[i]);
            }
            enqueue(e, act);
        } else if (work) {
            // Any schedule with u = 1 (kernels2d.cu:72-106,197-243): the tile
            // makes every attempt of the phase whose target lies in the tile
            // (with non-shared roles the target may be a sub-block's
            // neighbour site; the attempt then belongs to the target's tile,
            // which holds all the sites the attempt reads).  Candidates: the
            // active sub-blocks within c sites of the tile; the first NT
            // were staged in the previous phase (target and direction).
            const ResSched r = res_sched<LTE>(k, a.mcs0, p, x0, y0);
            const int ncand = r.ncx * r.ncy;
            for (int c0 = 0; c0 < ncand; c0 += NT) {
                const int ci = c0 + tid;
                ResEntry e[2];
                bool act[2] = {false, false};
                if (ci < ncand) {
                    uint2 t;
                    if (ci < NT) {
                        const uint4 d = sh.draws[ci];
                        t = make_uint2(d.x, d.y);
                    } else {
                        t = res_candidate_target<LTE>(k, r, ci, Lm, x0, y0);
                    }
                    if (t.x >> 31)
                        stage((int)(t.x & 0xFFu), (int)((t.x >> 8) & 0xFFu), (int)((t.x >> 16) & 0xFu), t.y, e[0],
                              act[0]);
                }
                enqueue(e, act);
            }
        }
        __syncthreads();
        // this phase's queue length, and whether the previous phase stopped
        // the launch (its result is then discarded by the host, so the
        // commit and publish that followed the stop did no harm: every
        // index they used stays in bounds)
        const uint2 cq = sh.ctl[p & 1u];
        // this thread's first queue entry, read with the length (entry tid
        // lies within the queue's storage; it is used only below the length)
        const ResEntry q0 = sh.queue.get(tid);
        if (cq.y) {
            if (tid == 0 && (cq.y & 1u)) atomicOr(a.bad, kResGuardMiss);
            return;
        }
        const int nqp = (int)cq.x;
        // The neighbours' packets of phase p - 1 were sent about when this
        // tile sent its own: loads issued now, while the queue is prepared,
        // usually find them (issued at the phase start they come back stale
        // and cost a second round trip).
#pragma unroll
        for (int jj = 0; jj < WPT; jj++)
            if (p > 0 && (need >> jj & 1u))
                v[jj] = __hip_atomic_load(a.mail + wsrc[jj] + par_in, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (gen) {
            // the previous phase's own deltas (sent with its packets) join
            // the rows, and their buffer is cleared for the next phase, by
            // the highest threads (idle here unless the queue is long); the
            // rows are next read after the second barrier.  (With c = 1 and
            // shared roles they join right after the publish: measured the
            // faster there, this the faster for the other schedules.)
            for (int s = NT - 1 - tid; s < nslots; s += NT) {
                const int64_t dv = sh.D[(p + 1) & 1u][s];
                if (dv != 0) {
                    atomicAdd(reinterpret_cast<unsigned long long*>(&sh.R[s]), (unsigned long long)dv);
                    sh.D[(p + 1) & 1u][s] = 0;
                }
            }
        }
        if (work) {
            if (tid == 0) sh.ctl[(p + 1) & 1u].x = 0;  // (last read in phase p - 1)
            // the previous phase's deltas are sent and in the rows
            if (!gen)
                for (int s = tid; s < nslots; s += NT) sh.D[(p + 1) & 1u][s] = 0;
            // prepare the interior attempts while the packets travel; a
            // thread's first attempt stays in its registers for the commit
            for (int qi = tid; qi < nqp; qi += NT) {
                // (every field in one round of loads, the ring test after)
                ResEntry qe = q0;
                if (qi != tid) qe = sh.queue.get(qi);
                asm volatile("" ::"v"(qe.li), "v"(qe.vs), "v"(qe.vt), "v"(qe.w2), "v"(qe.meta));
                const bool ring = (qe.li & kResRing) != 0;
                if (EARLY || !ring) {
                    bool same = false;
                    if (ring) {
                        // the ring's attempts too, on the halo as it stands
                        // (the commit redoes one whose 3 x 3 block holds a
                        // halo site that the packets change)
                        const int li = (int)(qe.li & ~kResRing);
                        int ox, oy, oz;
                        block_offset<2>((int)(qe.meta & 0xFu), ox, oy, oz);
                        qe.vt = sh.tile[li];
                        qe.vs = sh.tile[li + oy * S + ox];
                        same = qe.vs == qe.vt;  // (nothing to commit unless redone)
                        if (same) qe.meta &= 0xFu;
                    }
                    if (!same) {
                        if constexpr (LEAN)
                            res_prepare_lean<ST, LTE>(sh.tile, map, sh.p4, k.idbits, k.idmask, qe);
                        else
                            res_prepare<ST, LTE, XT>(lat, k, P, x0, y0, qe);
                    }
                    if (qi != tid) {
                        sh.queue.vs[qi] = qe.vs;
                        sh.queue.vt[qi] = qe.vt;
                        sh.queue.meta[qi] = qe.meta;
                        sh.queue.e[qi] = qe.e;
                    }
                }
                if (qi == tid) own = qe;
            }
        }
        if (p > 0) {
            long polls = 0;
            for (;;) {
#pragma unroll
                for (int jj = 0; jj < WPT; jj++)
                    if (((need & ~got) >> jj & 1u) && (uint32_t)(v[jj] >> 32) == p) got |= 1u << jj;
                if (got == need) break;
                if ((++polls & 63) == 0 &&
                    (__hip_atomic_load(a.bad, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u ||
                     __hip_atomic_load(a.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u))
                    break;
                if (polls > kBarrierSpinCap) {
                    __hip_atomic_store(a.err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
#pragma unroll
                for (int jj = 0; jj < WPT; jj++)
                    if ((need & ~got) >> jj & 1u)
                        v[jj] = __hip_atomic_load(a.mail + wsrc[jj] + par_in, __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT);
            }
            if (got != need) {
                *stop = 2u;
            } else {
#pragma unroll
                for (int jj = 0; jj < WPT; jj++) {
                    if (!(need >> jj & 1u)) continue;
                    const uint32_t x = win[jj], pl = (uint32_t)v[jj];
                    if (x >> 31) {
                        const int32_t pk = (int32_t)pl;
                        const int32_t dA = ((pk & 0xFFF) ^ 0x800) - 0x800;
                        const int64_t dP = (int64_t)((pk - dA) >> 12);
                        atomicAdd(reinterpret_cast<unsigned long long*>(&sh.R[x & 0xFFu]),
                                  (unsigned long long)(dP * 4294967296LL + dA));
                    } else {
                        const int li = (int)(x & 0xFFFFu), stride = (int)((x >> 16) & 0xFFFu);
                        sh.tile[li] = (ST)(SPW == 2 ? pl & 0xFFFFu : pl);
                        if (SPW == 2 && (x >> 28) > 1u) sh.tile[li + stride] = (ST)(pl >> 16);
                        if constexpr (EARLY) {
                            const uint32_t ch = pl ^ prev[jj];
                            sh.hdirty[li] = (uint8_t)((SPW == 2 ? ch & 0xFFFFu : ch) != 0u);
                            if (SPW == 2 && (x >> 28) > 1u) sh.hdirty[li + stride] = (uint8_t)((ch >> 16) != 0u);
                            prev[jj] = pl;
                        }
                    }
                }
            }
        }
        __syncthreads();
        if (!work) break;
        // the ring's attempts (their blocks reach the halo), then every
        // queued attempt's rows, acceptance and commit
        // The upper half of the threads, idle here unless the queue is long,
        // makes the next phase's draws: two pairs of sites each (one with a
        // thread per attempt), or the candidates' staged attempts.
        if (tid >= NT / 2 && p + 1 < Gp) {
            if (!gen) {
                for (int q = tid - NT / 2; q < NPAIR; q += NT / 2)
                    sh.draws[q] = res_pair_draws<LTE>(k, x0, y0, a.mcs0, p + 1, q);
            } else {
                const ResSched r = res_sched<LTE>(k, a.mcs0, p + 1, x0, y0);
                const int n = min(r.ncx * r.ncy, NT);
                for (int ci = tid - NT / 2; ci < n; ci += NT / 2) {
                    const uint2 t = res_candidate_target<LTE>(k, r, ci, Lm, x0, y0);
                    sh.draws[ci] = make_uint4(t.x, t.y, 0u, 0u);
                }
            }
        }
        const TileTables T{map, sh.R, sh.D[p & 1u], stop};
        bool miss = false;
        for (int qi = tid; qi < nqp; qi += NT) {
            ResEntry qe = own;
            if (qi != tid) {
                qe = sh.queue.get(qi);
                if constexpr (LEAN)
                    asm volatile("" ::"v"(qe.li), "v"(qe.vs), "v"(qe.vt), "v"(qe.w2), "v"(qe.meta), "v"(qe.e));
            }
            if (qe.li & kResRing) {
                const uint32_t li = qe.li & ~kResRing;
                bool redo = true;
                if constexpr (EARLY) {  // prepared before the packets: redone if its halo changed
                    const uint8_t* hd = sh.hdirty + li;
                    redo = (hd[-S - 1] | hd[-S] | hd[-S + 1] | hd[-1] | hd[1] | hd[S - 1] | hd[S] | hd[S + 1]) != 0;
                }
                qe.li = li;
                if (redo) {
                    int ox, oy, oz;
                    block_offset<2>((int)(qe.meta & 0xFu), ox, oy, oz);
                    qe.vt = sh.tile[li];
                    qe.vs = sh.tile[(int)li + oy * S + ox];
                    if (qe.vs == qe.vt) continue;  // kernels2d.cu:276
                    if constexpr (LEAN)
                        res_prepare_lean<ST, LTE>(sh.tile, map, sh.p4, k.idbits, k.idmask, qe);
                    else
                        res_prepare<ST, LTE, XT>(lat, k, P, x0, y0, qe);
                }
            }
            if constexpr (LEAN)
                res_commit_lean<ST, LTE>(sh.tile, sh.R, sh.D[p & 1u], sh.thr, k.nthr, sh.p4, k.idbits,
                                         k.idmask, qe, miss);
            else
                res_commit<ST, LTE>(sh.tile, k, P, T, sh.thr, thr_lds, qe);
        }
        if (miss) *stop = 1u;
        __syncthreads();
        // this tile's packets of phase p, tag p + 1
        const uint32_t par_out = (p & 1u) * 8u * mstride;
#pragma unroll
        for (int jj = 0; jj < WPT; jj++) {
            if (!(need >> jj & 1u)) continue;
            const uint32_t x = wout[jj];
            uint32_t pl;
            if (x >> 31) {
                const int64_t dv = sh.D[p & 1u][x & 0xFFu];
                const int32_t dA = (int32_t)(uint32_t)(uint64_t)dv;
                const int32_t dP = (int32_t)((dv - dA) >> 32);
                pl = (uint32_t)(dP * 4096 + dA);
            } else {
                const int li = (int)(x & 0xFFFFu), stride = (int)((x >> 16) & 0xFFFu);
                pl = (uint32_t)sh.tile[li];
                if (SPW == 2 && (x >> 28) > 1u) pl |= (uint32_t)sh.tile[li + stride] << 16;
            }
            __hip_atomic_store(a.mail + wdst[jj] + par_out, ((unsigned long long)(p + 1) << 32) | pl,
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        // the tile's own deltas join the rows (the neighbours' deltas of
        // this phase arrive with atomics too, possibly at the same time)
        if (!gen)
            for (int s = tid; s < nslots; s += NT) {
                const int64_t dv = sh.D[p & 1u][s];
                if (dv != 0) atomicAdd(reinterpret_cast<unsigned long long*>(&sh.R[s]), (unsigned long long)dv);
            }
    }
    if (sh.ctl[(Gp + 1) & 1u].y) return;  // the last wait gave up
    // the tile's sites and the rows of its cells (every holder has the same)
    for (int i = tid; i < TE * TE; i += NT) {
        const int ly = i >> LTE, lx = i & (TE - 1);
        glat[(uint32_t)(x0 + lx) | ((uint32_t)(y0 + ly) << k.lgL)] = sh.tile[(ly + 1) * S + lx + 1];
    }
    for (int s = tid; s < nslots; s += NT) {
        k.tab[0][sh.cell[s]] = sh.R[s];
        k.tab[1][sh.cell[s]] = sh.R[s];
    }
}

// VAR 0: the lean attempt path for c = 1 with shared roles; 3: the lean
// attempt path for any other schedule; 1: the general one; 2: the general
// one with the chemotaxis term.
// (one workgroup per CU: the compiler may schedule for four waves per SIMD
// rather than the most it could fit)
template <class ST, int LTE, int VAR>
__global__ __launch_bounds__((res_threads<LTE, VAR>())) __attribute__((amdgpu_waves_per_eu(1, 4))) void k_resident2d(
    ResidentArgs a) {
    constexpr int NT = res_threads<LTE, VAR>();
    extern __shared__ __align__(16) unsigned char res_map[];
    __shared__ ResShared<ST, LTE, NT> sh;
    constexpr bool LEAN = VAR == 0 || VAR == 3;
    if (LEAN || a.k.nt <= kSmallTypes) {
        SmallParams P;
        P.load(a.k.prm);
        resident_run<ST, LTE, NT, LEAN, VAR == 2, VAR == 0 ? 0 : VAR == 3 ? 1 : -1>(a, P, sh, res_map);
    } else {
        const int npar = param_words(a.k.nt);
        for (int i = (int)threadIdx.x; i < npar; i += NT) sh.prm[i] = a.k.prm->w[i];
        __syncthreads();
        const LdsParamView P{sh.prm, a.k.nt, a.k.nt * a.k.nt};
        resident_run<ST, LTE, NT, false, VAR == 2, -1>(a, P, sh, res_map);
    }
}

// Per-cell boxes of the lattice's sites (relative to the anchors, wrapped):
// bbox must hold (INT_MAX, INT_MIN, INT_MAX, INT_MIN) per cell on entry.
// A wave reads 64 sites of one row (L >= 64); each run of equal sites in
// it adds its first and last x and its row to the cell's box, from the
// run's first lane.  (A run whose relative x wraps belongs to a cell at
// least half the lattice wide: its box comes out wide either way.)
template <class ST>
__global__ __launch_bounds__(256) void k_bbox2d(const ST* __restrict__ lat, int lgL, uint32_t idmask,
                                                const int32_t* __restrict__ anchor, int32_t* bbox) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;  // (the grid covers the lattice exactly)
    const uint32_t c = (uint32_t)lat[i] & idmask;
    const int lane = (int)__lane_id();
    const uint32_t cl = (uint32_t)__shfl_up((int)c, 1), cr = (uint32_t)__shfl_down((int)c, 1);
    const bool first = lane == 0 || cl != c, last = lane == 63 || cr != c;
    const uint64_t lasts = __ballot(last);
    if (!first || c == 0u) return;
    const int end = lane + __ffsll((unsigned long long)(lasts >> lane)) - 1;  // this run's last lane
    const int L = 1 << lgL, Lm = L - 1;
    const int x0 = (int)(i & Lm), y = (int)(i >> lgL);
    const int ax = anchor[3 * c], ay = anchor[3 * c + 1];
    const int r0 = ((x0 - ax + L / 2) & Lm) - L / 2, r1 = ((x0 + end - lane - ax + L / 2) & Lm) - L / 2;
    const int ry = ((y - ay + L / 2) & Lm) - L / 2;
    atomicMin(bbox + 4 * c, min(r0, r1));
    atomicMax(bbox + 4 * c + 1, max(r0, r1));
    atomicMin(bbox + 4 * c + 2, ry);
    atomicMax(bbox + 4 * c + 3, ry);
}

// Holder words: the tiles whose sites or halo a cell's box overlaps, the box
// widened per axis by the largest margin up to res_margin that keeps it
// within two tiles (res_holds).  A cell whose box with its halo alone spans
// more than two tiles on an axis sets *bad; a cell with no site is held by no
// tile.
__device__ __forceinline__ bool res_axis(int lo, int hi, int lte, int mmax, int& t0, int& wide) {
    for (int m = mmax; m >= 0; m--) {
        const int a = (lo - 1 - m) >> lte, b = (hi + 1 + m) >> lte;  // (floor)
        if (b - a <= 1) {
            t0 = a;
            wide = b - a;
            return true;
        }
    }
    return false;
}

__global__ void k_res_holders(const int32_t* __restrict__ bbox, const int32_t* __restrict__ anchor, int rows,
                              int lte, int lnt, uint32_t* holders, unsigned int* bad) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= rows) return;
    const int32_t* q = bbox + 4 * c;
    if (c == 0 || q[1] < q[0]) {
        holders[c] = 0u;
        return;
    }
    const int mmax = res_margin(1 << lte);
    int tx0, wx, ty0, wy;
    if (!res_axis(anchor[3 * c] + q[0], anchor[3 * c] + q[1], lte, mmax, tx0, wx) ||
        !res_axis(anchor[3 * c + 1] + q[2], anchor[3 * c + 1] + q[3], lte, mmax, ty0, wy)) {
        atomicOr(bad, kResGuardWide);
        holders[c] = 0u;
        return;
    }
    const int ntm = (1 << lnt) - 1;
    holders[c] = (1u << 31) | (uint32_t)(tx0 & ntm) | ((uint32_t)(ty0 & ntm) << 12) | ((uint32_t)wx << 24) |
                 ((uint32_t)wy << 25);
}

template <class ST, int LTE>
static const void* resident_kernel(int var) {
    switch (var) {
        case 0: return (const void*)&k_resident2d<ST, LTE, 0>;
        case 1: return (const void*)&k_resident2d<ST, LTE, 1>;
        case 2: return (const void*)&k_resident2d<ST, LTE, 2>;
        default: return (const void*)&k_resident2d<ST, LTE, 3>;
    }
}

static const void* resident_kernel(bool narrow, int lte, int var) {
    if (narrow) return lte == 6 ? resident_kernel<uint16_t, 6>(var) : resident_kernel<uint16_t, 5>(var);
    return lte == 6 ? resident_kernel<uint32_t, 6>(var) : resident_kernel<uint32_t, 5>(var);
}

size_t resident_lds_bytes(int ncells) { return ((size_t)ncells + 1 + 15) & ~(size_t)15; }

int resident_threads(int lte, int var) {
    if (var == 0 || var == 3) return lte == 6 ? res_threads<6, 0>() : res_threads<5, 0>();
    return lte == 6 ? res_threads<6, 1>() : res_threads<5, 1>();
}

hipError_t resident_max_blocks(bool narrow, int lte, int var, int ncells, int* max_blocks) {
    int dev = 0, cus = 0, per_cu = 0;
    hipError_t e;
    if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
    if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
    if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, resident_kernel(narrow, lte, var),
                                                          resident_threads(lte, var), resident_lds_bytes(ncells))) !=
        hipSuccess)
        return e;
    *max_blocks = per_cu * cus;
    return hipSuccess;
}

hipError_t launch_resident(bool narrow, int var, const ResidentArgs& a, int blocks, hipStream_t s) {
    ResidentArgs copy = a;
    void* args[] = {&copy};
    return hipLaunchCooperativeKernel(resident_kernel(narrow, a.lte, var), dim3(blocks),
                                      dim3(resident_threads(a.lte, var)), args,
                                      (unsigned int)resident_lds_bytes(a.k.ncells), s);
}

// Everything a resident launch needs before its guard, in one kernel (the
// host issues one launch instead of six copies and fills): the lattice and
// the rows backed up, the counters' copy, the mailboxes and the error words
// zeroed, the boxes set empty.  Unit i of the grid: a 16-byte piece of the
// lattice copy, then of the mailboxes, then one row (copy and box), then the
// error words and counters.
__global__ __launch_bounds__(256) void k_res_prep(ResPrep q) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    const size_t u1 = q.lat_bytes / 16, u2 = q.nmail / 2, u3 = (size_t)q.rows;
    if (i < u1) {
        reinterpret_cast<uint4*>(q.lat_backup)[i] = reinterpret_cast<const uint4*>(q.lat)[i];
    } else if (i < u1 + u2) {
        reinterpret_cast<uint4*>(q.mail)[i - u1] = make_uint4(0u, 0u, 0u, 0u);
    } else if (i < u1 + u2 + u3) {
        const size_t c = i - u1 - u2;
        q.tab_backup[c] = q.tab[c];
        reinterpret_cast<int4*>(q.bbox)[c] = make_int4(INT_MAX, INT_MIN, INT_MAX, INT_MIN);
    } else if (i == u1 + u2 + u3) {
        q.flags[0] = q.flags[1] = 0u;
        if (q.counters)
            for (int j = 0; j < 4; j++) q.counters[4 + j] = q.counters[j];
    }
}

hipError_t launch_res_prep(const ResPrep& q, hipStream_t s) {
    if (q.lat_bytes % 16 != 0 || q.nmail % 2 != 0) return hipErrorInvalidValue;
    const size_t units = q.lat_bytes / 16 + q.nmail / 2 + (size_t)q.rows + 1;
    hipLaunchKernelGGL(k_res_prep, dim3((unsigned)((units + 255) / 256)), dim3(256), 0, s, q);
    return hipGetLastError();
}

hipError_t launch_resident_guard(bool narrow, const SweepConsts& k, const int32_t* anchor, int32_t* bbox,
                                 int lte, int lnt, uint32_t* holders, unsigned int* bad, hipStream_t s) {
    const int rows = k.ncells + 1;
    if (k.lgL < 7) return hipErrorInvalidValue;  // (the resident form's lattices: 128^2 and up)
    const uint32_t nsites = 1u << (2 * k.lgL);
    if (narrow)
        hipLaunchKernelGGL(k_bbox2d<uint16_t>, dim3((nsites + 255) / 256), dim3(256), 0, s, k.lat16, k.lgL,
                           k.idmask, anchor, bbox);
    else
        hipLaunchKernelGGL(k_bbox2d<uint32_t>, dim3((nsites + 255) / 256), dim3(256), 0, s, k.lat, k.lgL,
                           k.idmask, anchor, bbox);
    hipLaunchKernelGGL(k_res_holders, dim3((rows + 255) / 256), dim3(256), 0, s, bbox, anchor, rows, lte, lnt,
                       holders, bad);
    return hipGetLastError();
}

// Wave-aggregated updates of per-cell rows: the lanes of a wave are grouped
// by cell (the first pending lane's cell, then the next, ...), each group's
// contributions are summed across the wave and its first lane makes one
// atomic.  Neighbouring sites mostly share a cell, so a wave of 64 sites
// makes a handful of atomics instead of 64.  Every lane of the wave must call
// it (the loop and the shuffles are wave-wide).
__device__ __forceinline__ int64_t wave_sum(int64_t x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}
__device__ __forceinline__ uint32_t wave_min(uint32_t x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x = min(x, (uint32_t)__shfl_xor(x, o));
    return x;
}

// Area and perimeter recount (kernels2d.cu:5-41): one thread per site, the
// per-cell adds aggregated per wave.
template <int D>
__global__ __launch_bounds__(256) void k_tally_area_perim(const uint32_t* __restrict__ lat, int lgL,
                                                          int64_t* tab, uint32_t* first) {
    const uint32_t nsites = 1u << (D * lgL);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t c = i < nsites ? lat[i] & kSpinMask : 0u;
    int64_t add = 0;
    if (c) {
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
        add = (int64_t)cnt * 4294967296LL + 1;
    }
    const int lane = (int)__lane_id();
    uint64_t pending = __ballot(c != 0u);
    while (pending) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const uint32_t lc = (uint32_t)__shfl((int)c, leader);
        const bool mine = ((pending >> lane) & 1ull) && c == lc;
        const int64_t sum = wave_sum(mine ? add : 0);
        const uint32_t fmin = wave_min(mine ? i : 0xFFFFFFFFu);
        if (lane == leader) {
            atomic_add_i64(tab + lc, sum);
            atomicMin(first + lc, fmin);
        }
        pending &= ~__ballot(mine);
    }
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
    const uint32_t c = i < nsites ? lat[i] & kSpinMask : 0u;
    const int L = 1 << lgL, Lm = L - 1;
    const int pos[3] = {(int)(i & Lm), (int)((i >> lgL) & Lm), D == 3 ? (int)(i >> (2 * lgL)) : 0};
    int64_t w[3] = {0, 0, 0};
    if (c) {
#pragma unroll
        for (int ax = 0; ax < D; ax++) {
            const int v = pos[ax] - anchor[3 * c + ax];
            w[ax] = (int64_t)(((v + L / 2) & Lm) - L / 2);
        }
    }
    // per-wave sums by cell (see k_tally_area_perim)
    const int lane = (int)__lane_id();
    uint64_t pending = __ballot(c != 0u);
    while (pending) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const uint32_t lc = (uint32_t)__shfl((int)c, leader);
        const bool mine = ((pending >> lane) & 1ull) && c == lc;
#pragma unroll
        for (int ax = 0; ax < D; ax++) {
            const int64_t sum = wave_sum(mine ? w[ax] : 0);
            if (lane == leader) atomic_add_i64(csum + 3 * lc + ax, sum);
        }
        pending &= ~__ballot(mine);
    }
}

// 32-bit <-> 16-bit lattice (cell | type << 24 <-> cell | type << idbits),
// four sites per thread.
__global__ __launch_bounds__(256) void k_to_narrow(const uint4* __restrict__ w, uint2* n, uint32_t nq,
                                                   uint32_t idbits) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    const uint4 v = w[i];
    auto f = [idbits](uint32_t x) { return (x & kSpinMask) | ((x >> 24) << idbits); };
    n[i] = make_uint2(f(v.x) | (f(v.y) << 16), f(v.z) | (f(v.w) << 16));
}

__global__ __launch_bounds__(256) void k_to_wide(const uint2* __restrict__ n, uint4* w, uint32_t nq,
                                                 uint32_t idbits) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    const uint2 v = n[i];
    const uint32_t m = (1u << idbits) - 1u;
    auto f = [idbits, m](uint32_t x) { return (x & m) | ((x >> idbits) << 24); };
    w[i] = make_uint4(f(v.x & 0xFFFFu), f(v.x >> 16), f(v.y & 0xFFFFu), f(v.y >> 16));
}

hipError_t launch_convert(uint32_t* wide, uint16_t* narrow, size_t sites, int idbits, bool to_narrow,
                          hipStream_t s) {
    if (sites % 4 != 0 || sites / 4 > 0xFFFFFFFFu) return hipErrorInvalidValue;
    const uint32_t nq = (uint32_t)(sites / 4);
    const uint32_t blocks = (nq + 255) / 256;
    if (to_narrow)
        hipLaunchKernelGGL(k_to_narrow, dim3(blocks), dim3(256), 0, s, reinterpret_cast<const uint4*>(wide),
                           reinterpret_cast<uint2*>(narrow), nq, (uint32_t)idbits);
    else
        hipLaunchKernelGGL(k_to_wide, dim3(blocks), dim3(256), 0, s, reinterpret_cast<const uint2*>(narrow),
                           reinterpret_cast<uint4*>(wide), nq, (uint32_t)idbits);
    return hipGetLastError();
}

// Settle after G phases: fold the last phase's deltas into the lagging
// base, leaving both bases current and both delta buffers zero.
__global__ void k_settle(int64_t* tab, int64_t* del, int ncells) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c > ncells) return;
    tab[c] += del[c];
    del[c] = 0;
}

// Sufficient statistics of H(sigma; theta) (the backward pass of dH/dtheta)
// of many engines in one launch.  Block b serves engine r = eng_of[b]:
// blocks [pair0, pair0 + npair) count the unordered unlike-cell neighbour
// pairs by type pair; blocks [cell0, cell0 + ncell) sum the per-type cell
// moments.  Each unordered pair is seen once, from the site whose offset to
// the other is lexicographically positive ((dz, dy, dx) > 0: 4 offsets in
// 2D, 13 in 3D).
//
// Pairs, at most four types present: for each site every lane counts its
// unlike positive-half neighbours by neighbour type in a byte histogram (one
// byte per type of one 32-bit word) and adds it to the accumulator of its
// own site's type: acc[t] byte tn counts the (site type t, neighbour type tn)
// unlike pairs, at most 16 x 4 or 2 x 13 per lane.  The accumulators are
// widened to 16-bit fields and summed across the wave, and lane 0 adds
// M[a][b] + M[b][a] (M[a][a] on the diagonal) to the block's LDS counters;
// the block then adds them to the engine's global counters once.  More
// types: one LDS histogram add per unlike pair.
// Cells: a cell's type is fixed (flips copy whole values), so it is taken
// from the table made at push (ctype); per-type moments n, sum A, sum A^2,
// sum P, sum P^2 over live cells are summed across the wave, then added to
// the engine's counters (four types or fewer; else per-cell LDS adds).
// sites per thread in the pair blocks (in groups whose loads are all in
// flight at once: 16 x 5 values in 2D, 2 x 14 in 3D)
template <int D>
constexpr int stats_sites() { return D == 3 ? 2 : 32; }

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}

// Fold the lanes' byte accumulators acc[t] (byte tn: (site type t,
// neighbour type tn) unlike pairs, at most 64 per lane) into the block's LDS
// pair counters: bytes 0, 2 and bytes 1, 3 widened to 16-bit fields, summed
// across the wave; lane 0 adds M[a][b] + M[b][a] (M[a][a] on the diagonal).
__device__ __forceinline__ void stats_fold_four(const uint32_t acc[4], unsigned int* hist) {
    uint32_t w[8];
#pragma unroll
    for (int ty = 0; ty < 4; ty++) {
        w[2 * ty] = wave_sum_u32(acc[ty] & 0x00FF00FFu);
        w[2 * ty + 1] = wave_sum_u32((acc[ty] >> 8) & 0x00FF00FFu);
    }
    if (__lane_id() == 0) {
        auto M = [&](uint32_t t, uint32_t tn) {
            return (w[2 * t + (tn & 1u)] >> (16u * (tn >> 1))) & 0xFFFFu;
        };
#pragma unroll
        for (uint32_t a = 0; a < 4; a++)
#pragma unroll
            for (uint32_t b = a; b < 4; b++) {
                const uint32_t c = a == b ? M(a, a) : M(a, b) + M(b, a);
                if (c) atomicAdd(&hist[a * kNTypes + b], c);
            }
    }
}

// Sum of x over the wave, valid in every lane: DPP butterflies within each
// row of 16 lanes (quad_perm [1,0,3,2], [2,3,0,1], row_half_mirror,
// row_mirror leave every lane of a row holding the row's sum), then the
// four row sums read from lanes 15, 31, 47 and 63.
__device__ __forceinline__ uint32_t wave_sum_dpp(uint32_t x) {
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0xB1, 0xF, 0xF, false);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x4E, 0xF, 0xF, false);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x141, 0xF, 0xF, false);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x140, 0xF, 0xF, false);
    return (uint32_t)__builtin_amdgcn_readlane((int)x, 15) + (uint32_t)__builtin_amdgcn_readlane((int)x, 31) +
           (uint32_t)__builtin_amdgcn_readlane((int)x, 47) + (uint32_t)__builtin_amdgcn_readlane((int)x, 63);
}

// One chunk of 8 consecutive sites of a 2D 16-bit lattice (L >= 8) with
// every positive-half neighbour ((+1, 0), (-1, +1), (0, +1), (+1, +1)): one
// 16-byte load of the row, one of the row below, and three single sites at
// the chunk's ends.  Chunk c starts at site 8 c.
struct Chunk2d16 {
    uint4 a, b;          // row y: x0 .. x0 + 7; row y + 1: x0 .. x0 + 7
    uint32_t a8, bm, b8;  // row y: x0 + 8; row y + 1: x0 - 1, x0 + 8
};

__device__ __forceinline__ Chunk2d16 load_chunk_2d16(const uint16_t* __restrict__ lat, int lgL, uint32_t c) {
    const int Lm = (1 << lgL) - 1;
    const int x0 = (int)((c << 3) & (uint32_t)Lm), y = (int)((c << 3) >> lgL);
    const int y1 = (y + 1) & Lm;
    Chunk2d16 k;
    k.a = *reinterpret_cast<const uint4*>(lat + ((uint32_t)y << lgL) + x0);
    k.b = *reinterpret_cast<const uint4*>(lat + ((uint32_t)y1 << lgL) + x0);
    k.a8 = lat[((uint32_t)y << lgL) + ((x0 + 8) & Lm)];
    k.bm = lat[((uint32_t)y1 << lgL) + ((x0 - 1) & Lm)];
    k.b8 = lat[((uint32_t)y1 << lgL) + ((x0 + 8) & Lm)];
    return k;
}

// Unpack a chunk: r0[0..8] = row y, x0 .. x0 + 8; r1[0..9] = row y + 1,
// x0 - 1 .. x0 + 8.
__device__ __forceinline__ void unpack_chunk_2d16(const Chunk2d16& k, uint32_t r0[9], uint32_t r1[10]) {
    const uint32_t av[4] = {k.a.x, k.a.y, k.a.z, k.a.w}, bv[4] = {k.b.x, k.b.y, k.b.z, k.b.w};
#pragma unroll
    for (int j = 0; j < 4; j++) {
        r0[2 * j] = av[j] & 0xFFFFu;
        r0[2 * j + 1] = av[j] >> 16;
        r1[2 * j + 1] = bv[j] & 0xFFFFu;
        r1[2 * j + 2] = bv[j] >> 16;
    }
    r0[8] = k.a8;
    r1[0] = k.bm;
    r1[9] = k.b8;
}

// The pair counts of the three-type form (every site's type 0, 1 or 2) for
// 2D 16-bit lattices with L >= 8: a thread takes stats_sites<2>() / 8
// chunks.  Per site, hh = sum over its unlike positive-half neighbours of
// 1 << 7 tn (three 7-bit slots, at most 4 each); the chunk's sum of
// hh << 21 t keeps the nine (t, tn) counts in 7-bit fields f = 3 t + tn of
// one 64-bit word (at most 32 each per chunk).  Two chunks are added (at
// most 64), then split into the even and odd fields, which have 14 bits of
// room apart, and accumulated.  At the end the fields are widened to 16-bit
// halves of five words, summed across the wave, and one lane adds the six
// unordered type-pair counts (M[a][b] + M[b][a], M[a][a] on the diagonal)
// to the block's LDS counters.
__device__ __forceinline__ void stats_pairs_2d16_t3(const StatsDesc& e, int lb, unsigned int* hist) {
    constexpr int NCH = stats_sites<2>() / 8;
    static_assert(NCH % 2 == 0, "chunks are added in pairs");
    constexpr uint64_t kEven = 0x7Full | 0x7Full << 14 | 0x7Full << 28 | 0x7Full << 42 | 0x7Full << 56;
    const int lgL = e.lgL;
    const uint32_t nchunks = 1u << (2 * lgL - 3);
    const uint32_t idb = e.idbits;
    const uint16_t* __restrict__ lat = static_cast<const uint16_t*>(e.lat);
    uint64_t E = 0, O = 0, two = 0;
    auto chunk_of = [&](int h) { return ((uint32_t)lb * NCH + (uint32_t)h) * kBlock + threadIdx.x; };
    Chunk2d16 k = load_chunk_2d16(lat, lgL, chunk_of(0) < nchunks ? chunk_of(0) : 0u);
#pragma unroll
    for (int h = 0; h < NCH; h++) {
        const bool ok = chunk_of(h) < nchunks;
        Chunk2d16 kn = k;
        if (h + 1 < NCH) kn = load_chunk_2d16(lat, lgL, chunk_of(h + 1) < nchunks ? chunk_of(h + 1) : 0u);
        uint32_t r0[9], r1[10];
        unpack_chunk_2d16(k, r0, r1);
        uint32_t o0[9], o1[10];  // 1 << 7 t of each neighbour
#pragma unroll
        for (int j = 1; j < 9; j++) o0[j] = 1u << ((r0[j] >> idb) * 7u);
#pragma unroll
        for (int j = 0; j < 10; j++) o1[j] = 1u << ((r1[j] >> idb) * 7u);
        uint64_t acc = 0;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const uint32_t v = r0[j];
            const uint32_t hh = (r0[j + 1] != v ? o0[j + 1] : 0u) + (r1[j] != v ? o1[j] : 0u) +
                                (r1[j + 1] != v ? o1[j + 1] : 0u) + (r1[j + 2] != v ? o1[j + 2] : 0u);
            acc += (uint64_t)hh << ((v >> idb) * 21u);
        }
        two += ok ? acc : 0ull;
        if (h & 1) {
            E += two & kEven;
            O += (two >> 7) & kEven;
            two = 0;
        }
        k = kn;
    }
    // fields f = 3 t + tn: even f = 2 i in E at 14 i, odd f = 2 i + 1 in O at 14 i
    auto fe = [&](int i) { return (uint32_t)(E >> (14 * i)) & 0x3FFFu; };
    auto fo = [&](int i) { return (uint32_t)(O >> (14 * i)) & 0x3FFFu; };
    uint32_t w[5] = {fe(0) | fo(0) << 16, fe(1) | fo(1) << 16, fe(2) | fo(2) << 16, fe(3) | fo(3) << 16, fe(4)};
#pragma unroll
    for (int q = 0; q < 5; q++) w[q] = wave_sum_dpp(w[q]);
    if (__lane_id() == 0) {
        auto f = [&](int i) { return (w[i >> 1] >> (16 * (i & 1))) & 0xFFFFu; };
        const uint32_t c00 = f(0), c01 = f(1) + f(3), c02 = f(2) + f(6), c11 = f(4), c12 = f(5) + f(7),
                       c22 = f(8);
        if (c00) atomicAdd(&hist[0 * kNTypes + 0], c00);
        if (c01) atomicAdd(&hist[0 * kNTypes + 1], c01);
        if (c02) atomicAdd(&hist[0 * kNTypes + 2], c02);
        if (c11) atomicAdd(&hist[1 * kNTypes + 1], c11);
        if (c12) atomicAdd(&hist[1 * kNTypes + 2], c12);
        if (c22) atomicAdd(&hist[2 * kNTypes + 2], c22);
    }
}

// The pair counts of the four-type form for 2D 16-bit lattices with L >= 8,
// read as chunks (Chunk2d16); a thread takes stats_sites<2>() / 8 chunks,
// the next chunk's loads issued before the current one is counted.
__device__ __forceinline__ void stats_pairs_2d16(const StatsDesc& e, int lb, unsigned int* hist) {
    constexpr int NCH = stats_sites<2>() / 8;
    const int lgL = e.lgL;
    const uint32_t nchunks = 1u << (2 * lgL - 3);
    const uint32_t idb = e.idbits;
    const uint16_t* __restrict__ lat = static_cast<const uint16_t*>(e.lat);
    uint32_t acc[4] = {0u, 0u, 0u, 0u};  // bytes: at most 4 x 32 per lane
    auto chunk_of = [&](int h) { return ((uint32_t)lb * NCH + (uint32_t)h) * kBlock + threadIdx.x; };
    Chunk2d16 k = load_chunk_2d16(lat, lgL, chunk_of(0) < nchunks ? chunk_of(0) : 0u);
#pragma unroll
    for (int h = 0; h < NCH; h++) {
        const bool ok = chunk_of(h) < nchunks;
        Chunk2d16 kn = k;
        if (h + 1 < NCH) kn = load_chunk_2d16(lat, lgL, chunk_of(h + 1) < nchunks ? chunk_of(h + 1) : 0u);
        uint32_t r0[9], r1[10];
        unpack_chunk_2d16(k, r0, r1);
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const uint32_t v = r0[j];
            const uint32_t nb[4] = {r0[j + 1], r1[j], r1[j + 1], r1[j + 2]};
            uint32_t hh = 0;
#pragma unroll
            for (int q = 0; q < 4; q++) hh += nb[q] != v ? 1u << ((nb[q] >> idb) << 3) : 0u;
            hh = ok ? hh : 0u;
            const uint32_t t = v >> idb;
#pragma unroll
            for (uint32_t ty = 0; ty < 4; ty++) acc[ty] += t == ty ? hh : 0u;
        }
        k = kn;
    }
    stats_fold_four(acc, hist);
}

// T3: every engine of the launch is 2D with 16-bit sites, L >= 8 and at
// most three types (the three-type chunk form, a kernel of its own so that
// its register budget is not the other forms').
template <int D, class ST, bool T3>
__device__ __forceinline__ void stats_pairs_block(const StatsDesc& e, int lb) {
    __shared__ unsigned int hist[kNTypes * kNTypes];
    constexpr int NB = D == 3 ? 27 : 9;
    constexpr int NH = NB / 2;
    constexpr int kStatsSites = stats_sites<D>();
    const int lgL = e.lgL, Lm = (1 << lgL) - 1;
    const uint32_t nsites = 1u << (D * lgL);
    const bool four = e.nt <= 4;
    for (int i = threadIdx.x; i < kNTypes * kNTypes; i += blockDim.x) hist[i] = 0;
    __syncthreads();
    if constexpr (T3) {
        static_assert(D == 2 && sizeof(ST) == 2, "the three-type chunk form is 2D, 16-bit");
        stats_pairs_2d16_t3(e, lb, hist);
        __syncthreads();
        for (int i = threadIdx.x; i < kNTypes * kNTypes; i += blockDim.x)
            if (hist[i]) atomicAdd(&e.out[i], (unsigned long long)hist[i]);
        return;
    }
    if constexpr (D == 2 && sizeof(ST) == 2) {
        if (four && lgL >= 3) {  // (uniform per block)
            stats_pairs_2d16(e, lb, hist);
            __syncthreads();
            for (int i = threadIdx.x; i < kNTypes * kNTypes; i += blockDim.x)
                if (hist[i]) atomicAdd(&e.out[i], (unsigned long long)hist[i]);
            return;
        }
    }
    const ST* __restrict__ lat = static_cast<const ST*>(e.lat);
    auto site = [&](uint32_t j) -> uint32_t { return (uint32_t)lat[j]; };
    constexpr int G = kStatsSites < 8 ? kStatsSites : 8;  // sites per group
    uint32_t acc[4] = {0u, 0u, 0u, 0u};  // bytes: at most 13 x 2 (3D) or 4 x 32 (2D) per lane
    for (int g0 = 0; g0 < kStatsSites; g0 += G) {
    const uint32_t base = (uint32_t)lb * (kBlock * kStatsSites) + (uint32_t)g0 * kBlock;
    // every load of a group's sites and their positive-half neighbours is
    // issued before any comparison (one memory round trip, not one per site)
    uint32_t vs[G], vn_[G][NH];
#pragma unroll
    for (int k = 0; k < G; k++) {
        const uint32_t i = base + (uint32_t)k * kBlock + threadIdx.x;
        const uint32_t ii = i < nsites ? i : 0u;
        const int x = (int)(ii & Lm), y = (int)((ii >> lgL) & Lm), z = D == 3 ? (int)(ii >> (2 * lgL)) : 0;
        vs[k] = site(ii);
#pragma unroll
        for (int n = NB / 2 + 1; n < NB; n++) {  // offsets after the centre: the positive half
            int ox, oy, oz;
            block_offset<D>(n, ox, oy, oz);
            vn_[k][n - NB / 2 - 1] = site(site_index<D>((x + ox) & Lm, (y + oy) & Lm, (z + oz) & Lm, lgL));
        }
    }
    if (four) {  // (uniform per block)
#pragma unroll
        for (int k = 0; k < G; k++) {
            const uint32_t i = base + (uint32_t)k * kBlock + threadIdx.x;
            const uint32_t v = vs[k];
            uint32_t h = 0;
#pragma unroll
            for (int q = 0; q < NH; q++) {
                const uint32_t vn = vn_[k][q];
                h += vn != v ? 1u << ((vn >> e.idbits) << 3) : 0u;
            }
            h = i < nsites ? h : 0u;
            const uint32_t t = v >> e.idbits;
#pragma unroll
            for (uint32_t ty = 0; ty < 4; ty++) acc[ty] += t == ty ? h : 0u;
        }
    } else {
#pragma unroll
        for (int k = 0; k < G; k++) {
            const uint32_t i = base + (uint32_t)k * kBlock + threadIdx.x;
            if (i >= nsites) continue;
            const uint32_t v = vs[k];
            const uint32_t t = v >> e.idbits;
#pragma unroll
            for (int q = 0; q < NH; q++) {
                const uint32_t vn = vn_[k][q];
                if (vn != v) {
                    const uint32_t tn = vn >> e.idbits;
                    atomicAdd(&hist[min(t, tn) * kNTypes + max(t, tn)], 1u);
                }
            }
        }
    }
    }
    if (four) stats_fold_four(acc, hist);
    __syncthreads();
    for (int i = threadIdx.x; i < kNTypes * kNTypes; i += blockDim.x)
        if (hist[i]) atomicAdd(&e.out[i], (unsigned long long)hist[i]);
}

// cells per thread in the cell blocks (summed in registers before the wave
// sums)
constexpr int kStatsCells = 8;

__device__ __forceinline__ void stats_cells_block(const StatsDesc& e, int lb) {
    __shared__ unsigned long long acc[kNTypes * 5];
    const bool four = e.nt <= 4;
    for (int i = threadIdx.x; i < kNTypes * 5; i += blockDim.x) acc[i] = 0;
    __syncthreads();
    if (four) {
        // the thread's cells (area, perimeter and type; area 0 for none)
        uint32_t A[kStatsCells], P[kStatsCells];
        int T[kStatsCells];
#pragma unroll
        for (int k = 0; k < kStatsCells; k++) {
            const int c = 1 + (lb * kStatsCells + k) * kBlock + (int)threadIdx.x;
            const bool in = c <= e.ncells;
            const int64_t row = in ? e.tab[c] : 0;
            A[k] = (uint32_t)row;
            P[k] = (uint32_t)(row >> 32);
            T[k] = in ? e.ctype[c] : 0;
        }
        for (int ty = 0; ty < e.nt; ty++) {  // (uniform: no cell has another type)
            uint64_t v[5] = {0, 0, 0, 0, 0};
#pragma unroll
            for (int k = 0; k < kStatsCells; k++) {
                const bool mine = A[k] != 0 && T[k] == ty;
                const uint64_t a = mine ? A[k] : 0u, p = mine ? P[k] : 0u;
                v[0] += mine ? 1u : 0u;
                v[1] += a;
                v[2] += a * a;
                v[3] += p;
                v[4] += p * p;
            }
#pragma unroll
            for (int m = 0; m < 5; m++) {
                const uint64_t sum = wave_sum_u64(v[m]);
                if (__lane_id() == 0 && sum) atomicAdd(&acc[5 * ty + m], (unsigned long long)sum);
            }
        }
    } else {
        for (int k = 0; k < kStatsCells; k++) {
            const int c = 1 + (lb * kStatsCells + k) * kBlock + (int)threadIdx.x;
            if (c > e.ncells) break;
            const int64_t row = e.tab[c];
            const uint64_t A = (uint32_t)row, P = (uint64_t)(row >> 32);
            if (!A) continue;
            unsigned long long* m = acc + 5 * e.ctype[c];
            atomicAdd(&m[0], 1ull);
            atomicAdd(&m[1], (unsigned long long)A);
            atomicAdd(&m[2], (unsigned long long)(A * A));
            atomicAdd(&m[3], (unsigned long long)P);
            atomicAdd(&m[4], (unsigned long long)(P * P));
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kNTypes * 5; i += blockDim.x)
        if (acc[i]) atomicAdd(&e.out[kNTypes * kNTypes + i], acc[i]);
}

// ST: the site type of every engine of the launch (16-bit or 32-bit sites)
template <int D, class ST, bool T3>
__global__ __launch_bounds__(kBlock) void k_stats_batch(const StatsDesc* __restrict__ descs,
                                                        const int* __restrict__ eng_of) {
    const int r = eng_of[blockIdx.x];
    const StatsDesc e = descs[r];  // uniform per block: scalar loads
    const int b = (int)blockIdx.x;
    if (b < e.cell0)
        stats_pairs_block<D, ST, T3>(e, b - e.pair0);
    else
        stats_cells_block(e, b - e.cell0);
}

// The allocation-free statistics launch (cpm_energy_stats_many_noalloc):
// up to kStatsArgEngines engines per launch, their descriptors passed by
// value as the kernel argument; block b serves the engine whose block range
// [pair0, end) holds it.
template <int D, class ST, bool T3>
__global__ __launch_bounds__(kBlock) void k_stats_args(StatsArgs a) {
    const int b = (int)blockIdx.x;
    int r = 0;
    while (r + 1 < a.n && b >= a.end[r]) r++;  // uniform per block
    const StatsDesc e = a.d[r];
    if (b < e.cell0)
        stats_pairs_block<D, ST, T3>(e, b - e.pair0);
    else
        stats_cells_block(e, b - e.cell0);
}

__global__ void k_fill_u64(unsigned long long* p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

// Scratch of the allocation-free statistics entry point: the statistics of
// kStatsScratchEngines engines (400 pair counts + 100 moments each).  Each
// GPU holds its own copy until the process ends; its address is looked up
// once per device (the array's name is not a device address on the host).
static __device__ __align__(256) unsigned long long scratch[kStatsScratchEngines * kStatsWords];

hipError_t stats_scratch(unsigned long long** out) {
    static std::mutex mu;
    static unsigned long long* ptr[kMaxDevices] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(mu);
    if (!ptr[dev]) {
        void* p = nullptr;
        if ((e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch))) != hipSuccess) return e;
        ptr[dev] = static_cast<unsigned long long*>(p);
    }
    *out = ptr[dev];
    return hipSuccess;
}

hipError_t launch_fill_zero(unsigned long long* p, size_t n, hipStream_t s) {
    const size_t blocks = std::min<size_t>((n + 255) / 256, 1024);
    hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)std::max<size_t>(blocks, 1)), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

static_assert(sizeof(StatsArgs) <= 4096, "the descriptors are passed as one kernel argument");

hipError_t launch_stats_args(int d, bool narrow, bool t3, const StatsArgs& a, int blocks, hipStream_t s) {
#define CPM_STATS(DD, ST, T3) hipLaunchKernelGGL((k_stats_args<DD, ST, T3>), dim3(blocks), dim3(kBlock), 0, s, a)
    if (d == 2) {
        if (narrow && t3) CPM_STATS(2, uint16_t, true);
        else if (narrow) CPM_STATS(2, uint16_t, false);
        else CPM_STATS(2, uint32_t, false);
    } else {
        if (narrow) CPM_STATS(3, uint16_t, false); else CPM_STATS(3, uint32_t, false);
    }
#undef CPM_STATS
    return hipGetLastError();
}

int stats_pair_blocks(int d, int lgL) {
    const long nsites = 1L << (d * lgL);
    const long per = (long)kBlock * (d == 3 ? stats_sites<3>() : stats_sites<2>());
    return (int)((nsites + per - 1) / per);
}

int stats_cell_blocks(int ncells) { return (ncells + kBlock * kStatsCells - 1) / (kBlock * kStatsCells); }

hipError_t launch_stats_batch(int d, bool narrow, bool t3, const StatsDesc* descs, const int* eng_of, int blocks,
                              hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
#define CPM_STATS(DD, ST, T3) \
    hipExtLaunchKernelGGL((k_stats_batch<DD, ST, T3>), dim3(blocks), dim3(kBlock), 0, s, e0, e1, 0, descs, eng_of)
    if (d == 2) {
        if (narrow && t3) CPM_STATS(2, uint16_t, true);
        else if (narrow) CPM_STATS(2, uint16_t, false);
        else CPM_STATS(2, uint32_t, false);
    } else {
        if (narrow) CPM_STATS(3, uint16_t, false); else CPM_STATS(3, uint32_t, false);
    }
#undef CPM_STATS
    return hipGetLastError();
}

// The per-phase kernel a launch uses (results never depend on the choice).
enum PhaseKernel {
    kPkBox3, kPkCompact2Small, kPkCompact2SmallFull, kPkCompact2Lds, kPkCompact2LdsFull,
    kPkCompact3Small, kPkCompact3Lds, kPkPhase2Small, kPkPhase2Lds, kPkPhase3Small, kPkPhase3Lds,
    kPkCompact2Lds2, kPkCompact2Lds4, kPkCompact3Small2, kPkCompact3Small4, kPkCompact3Small8,
    kPkCompact2Lds8, kPkRef2, kPkRef3, kPkCount
};

// Kernels on the 16-bit lattice have ids kPkNarrow + id; the 3D scalar
// compacted kernels without the extra terms kPkLean + that.
constexpr int kPkNarrow = 64;
constexpr int kPkLean = 128;

const char* phase_kernel_name(int id) {
    static const char* const base[kPkCount] = {
        "k_sweep_phase3_box",
        "k_sweep_phase_compact<2, false, false, 1, %s, true>",
        "k_sweep_phase_compact<2, false, true, 1, %s, true>",
        "k_sweep_phase_compact<2, true, false, 1, %s, true>",
        "k_sweep_phase_compact<2, true, true, 1, %s, true>",
        "k_sweep_phase_compact<3, false, false, 1, %s, true>",
        "k_sweep_phase_compact<3, true, false, 1, %s, true>",
        "k_sweep_phase<2, %s>", "k_sweep_phase_lds<2, %s>",
        "k_sweep_phase<3, %s>", "k_sweep_phase_lds<3, %s>",
        "k_sweep_phase_compact<2, true, false, 2, %s, true>",
        "k_sweep_phase_compact<2, true, false, 4, %s, true>",
        "k_sweep_phase_compact<3, false, false, 2, %s, true>",
        "k_sweep_phase_compact<3, false, false, 4, %s, %s>",
        "k_sweep_phase_compact<3, false, false, 8, %s, %s>",
        "k_sweep_phase_compact<2, true, false, 8, %s, true>",
        "k_sweep_phase_ref<2, %s>", "k_sweep_phase_ref<3, %s>"};
    static char names[4][kPkCount][96];
    static bool made = false;
    if (!made) {
        for (int v = 0; v < 4; v++)
            for (int i = 0; i < kPkCount; i++)
                snprintf(names[v][i], sizeof(names[v][i]), base[i],
                         (v & 1) ? "unsigned short" : "unsigned int", (v & 2) ? "false" : "true");
        made = true;
    }
    if (id < 0) return "";
    const int v = ((id & kPkNarrow) ? 1 : 0) | ((id & kPkLean) ? 2 : 0);
    id &= kPkNarrow - 1;
    if (id >= kPkCount || (id == kPkBox3 && v)) return "";
    return id == kPkBox3 ? base[kPkBox3] : names[v][id];
}

static int choose_phase_kernel(int d, const PhaseArgs& a, bool box_ok) {
    const int L = 1 << a.k.lgL;
    if (a.k.refstream) return d == 2 ? kPkRef2 : kPkRef3;
    if (box_ok && d == 3 && a.k.lc == 0 && !a.k.roles && a.lstride == 1 && L >= kBX && L >= kBY &&
        L >= kBZ && a.box3 && !a.sync_pass) {
        const long nbox = (long)(L / kBX) * (L / kBY) * (L / kBZ);
        if (nbox * kBlock >= a.k.ncells + 1) return kPkBox3;
    }
    const bool small = a.k.nt <= kSmallTypes && (a.form & kFormSmall);
    const bool compact = a.k.u == 1 && (a.form & kFormCompact) && a.sync_pass < 3;
    const bool full = a.form & kFormFull;
    const int att = form_att(a.form);
    if (compact) {
        if (d == 2 && !small && !full && att > 1)
            return att == 8 ? kPkCompact2Lds8 : att == 4 ? kPkCompact2Lds4 : kPkCompact2Lds2;
        if (d == 3 && small && att > 1)
            return att == 8 ? kPkCompact3Small8 : att == 4 ? kPkCompact3Small4 : kPkCompact3Small2;
        if (d == 2) return small ? (full ? kPkCompact2SmallFull : kPkCompact2Small)
                                 : (full ? kPkCompact2LdsFull : kPkCompact2Lds);
        return small ? kPkCompact3Small : kPkCompact3Lds;
    }
    if (d == 2) return small ? kPkPhase2Small : kPkPhase2Lds;
    return small ? kPkPhase3Small : kPkPhase3Lds;
}

template <class ST>
static void launch_phase_st(int id, const PhaseArgs& a, hipStream_t s, hipEvent_t e0, hipEvent_t e1,
                            bool lean) {
    const int work = a.nsub > a.k.ncells + 1 ? a.nsub : a.k.ncells + 1;
    const int blocks = (work + kBlock - 1) / kBlock;
    // hipExtLaunchKernelGGL records e0/e1 at the dispatch's start and end
    // (null events: plain launch).
#define CPM_LAUNCH(KERNEL, GRID) \
    hipExtLaunchKernelGGL(KERNEL, dim3(GRID), dim3(kBlock), 0, s, e0, e1, 0, a)
    // blocks of kBlock threads drawing att attempts each
    auto cblocks = [&](int att) {
        const int per = att * kBlock;
        const int nb = (a.nsub + per - 1) / per;
        return nb > 0 ? nb : 1;
    };
    switch (id) {
        case kPkCompact2Small: CPM_LAUNCH((k_sweep_phase_compact<2, false, false, 1, ST, true>), blocks); break;
        case kPkCompact2SmallFull: CPM_LAUNCH((k_sweep_phase_compact<2, false, true, 1, ST, true>), blocks); break;
        case kPkCompact2Lds: CPM_LAUNCH((k_sweep_phase_compact<2, true, false, 1, ST, true>), blocks); break;
        case kPkCompact2LdsFull: CPM_LAUNCH((k_sweep_phase_compact<2, true, true, 1, ST, true>), blocks); break;
        case kPkCompact3Small: CPM_LAUNCH((k_sweep_phase_compact<3, false, false, 1, ST, true>), blocks); break;
        case kPkCompact3Lds: CPM_LAUNCH((k_sweep_phase_compact<3, true, false, 1, ST, true>), blocks); break;
        case kPkCompact2Lds2: CPM_LAUNCH((k_sweep_phase_compact<2, true, false, 2, ST, true>), cblocks(2)); break;
        case kPkCompact2Lds4: CPM_LAUNCH((k_sweep_phase_compact<2, true, false, 4, ST, true>), cblocks(4)); break;
        case kPkCompact3Small2: CPM_LAUNCH((k_sweep_phase_compact<3, false, false, 2, ST, true>), cblocks(2)); break;
        case kPkCompact3Small4:
            if (lean) CPM_LAUNCH((k_sweep_phase_compact<3, false, false, 4, ST, false>), cblocks(4));
            else CPM_LAUNCH((k_sweep_phase_compact<3, false, false, 4, ST, true>), cblocks(4));
            break;
        case kPkCompact3Small8:
            if (lean) CPM_LAUNCH((k_sweep_phase_compact<3, false, false, 8, ST, false>), cblocks(8));
            else CPM_LAUNCH((k_sweep_phase_compact<3, false, false, 8, ST, true>), cblocks(8));
            break;
        case kPkCompact2Lds8: CPM_LAUNCH((k_sweep_phase_compact<2, true, false, 8, ST, true>), cblocks(8)); break;
        case kPkPhase2Small: CPM_LAUNCH((k_sweep_phase<2, ST>), blocks); break;
        case kPkPhase2Lds: CPM_LAUNCH((k_sweep_phase_lds<2, ST>), blocks); break;
        case kPkPhase3Small: CPM_LAUNCH((k_sweep_phase<3, ST>), blocks); break;
        case kPkRef2: CPM_LAUNCH((k_sweep_phase_ref<2, ST>), blocks); break;
        case kPkRef3: CPM_LAUNCH((k_sweep_phase_ref<3, ST>), blocks); break;
        default: CPM_LAUNCH((k_sweep_phase_lds<3, ST>), blocks); break;
    }
#undef CPM_LAUNCH
}

hipError_t launch_phase(int d, const PhaseArgs& a, hipStream_t s, hipEvent_t e0, hipEvent_t e1,
                        int* kernel_id) {
    const int L = 1 << a.k.lgL;
    const bool narrow = a.k.lat16 != nullptr;
    // the LDS box form reads the 32-bit lattice only
    const int id = choose_phase_kernel(d, a, !narrow);
    // without chemotaxis, act and persistence the 3D scalar kernels run lean
    const bool lean = !a.k.chem && !a.k.act && !a.k.pers &&
                      (id == kPkCompact3Small4 || id == kPkCompact3Small8);
    if (kernel_id) *kernel_id = id + (narrow ? kPkNarrow : 0) + (lean ? kPkLean : 0);
    if (id == kPkBox3)
        hipExtLaunchKernelGGL(k_sweep_phase3_box, dim3((L / kBX) * (L / kBY) * (L / kBZ)), dim3(kBlock),
                              0, s, e0, e1, 0, a);
    else if (narrow)
        launch_phase_st<uint16_t>(id, a, s, e0, e1, lean);
    else
        launch_phase_st<uint32_t>(id, a, s, e0, e1, lean);
    return hipGetLastError();
}

__global__ void k_set_u32(uint32_t* p, uint32_t v) { *p = v; }
__global__ void k_add_u32(uint32_t* p, uint32_t v) { *p += v; }

hipError_t launch_set_u32(uint32_t* p, uint32_t v, hipStream_t s) {
    hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, s, p, v);
    return hipGetLastError();
}

hipError_t launch_add_u32(uint32_t* p, uint32_t v, hipStream_t s) {
    hipLaunchKernelGGL(k_add_u32, dim3(1), dim3(1), 0, s, p, v);
    return hipGetLastError();
}

hipError_t batch_max_blocks(int d, int* max_blocks) {
    int dev = 0, cus = 0, per_cu = 0;
    hipError_t e;
    if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
    if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess)
        return e;
    const void* f = d == 2 ? reinterpret_cast<const void*>(&k_batch_global<2>)
                           : reinterpret_cast<const void*>(&k_batch_global<3>);
    if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, kBlock, 0)) != hipSuccess)
        return e;
    *max_blocks = per_cu * cus;
    return hipSuccess;
}

hipError_t launch_batch(int d, const BatchArgs& a, int blocks, hipStream_t s) {
    BatchArgs copy = a;
    void* args[] = {&copy};
    const void* f = d == 2 ? reinterpret_cast<const void*>(&k_batch_global<2>)
                           : reinterpret_cast<const void*>(&k_batch_global<3>);
    return hipLaunchCooperativeKernel(f, dim3(blocks), dim3(kBlock), args, 0, s);
}

template <class T>
static void launch_batch_phase_st(int d, const BatchPhaseArgs& a, int blocks, hipStream_t s,
                                  hipEvent_t e0, hipEvent_t e1, char* name, size_t name_len) {
    const bool compact = a.form & kFormCompact;
    const bool small = a.form & kFormSmall;
    const int att = form_att(a.form) >= 8 ? 8 : form_att(a.form) >= 4 ? 4 : 1;
    const char* st = sizeof(T) == 2 ? "unsigned short" : "unsigned int";
    // hipExtLaunchKernelGGL records e0/e1 at the dispatch's start and end
    // (null events: plain launch)
#define CPM_BATCH(DD, SG, AT) \
    hipExtLaunchKernelGGL((k_batch_phase_compact<DD, SG, AT, T>), dim3(blocks), dim3(kBlock), 0, s, e0, e1, 0, a)
    if (compact) {
        if (name) snprintf(name, name_len, "k_batch_phase_compact<%d, %s, %d, %s>", d, small ? "false" : "true", att, st);
        if (d == 2) {
            if (small) {
                if (att == 8) CPM_BATCH(2, false, 8); else if (att == 4) CPM_BATCH(2, false, 4); else CPM_BATCH(2, false, 1);
            } else {
                if (att == 8) CPM_BATCH(2, true, 8); else if (att == 4) CPM_BATCH(2, true, 4); else CPM_BATCH(2, true, 1);
            }
        } else {
            if (small) {
                if (att == 8) CPM_BATCH(3, false, 8); else if (att == 4) CPM_BATCH(3, false, 4); else CPM_BATCH(3, false, 1);
            } else {
                if (att == 8) CPM_BATCH(3, true, 8); else if (att == 4) CPM_BATCH(3, true, 4); else CPM_BATCH(3, true, 1);
            }
        }
        return;
    }
#undef CPM_BATCH
    if (name) snprintf(name, name_len, "k_batch_phase<%d, %s%s>", d, st, a.refstream ? ", true" : "");
    if (a.refstream) {
        if (d == 2)
            hipExtLaunchKernelGGL((k_batch_phase<2, T, true>), dim3(blocks), dim3(kBlock), 0, s, e0, e1, 0, a);
        else
            hipExtLaunchKernelGGL((k_batch_phase<3, T, true>), dim3(blocks), dim3(kBlock), 0, s, e0, e1, 0, a);
    } else if (d == 2) {
        hipExtLaunchKernelGGL((k_batch_phase<2, T>), dim3(blocks), dim3(kBlock), 0, s, e0, e1, 0, a);
    } else {
        hipExtLaunchKernelGGL((k_batch_phase<3, T>), dim3(blocks), dim3(kBlock), 0, s, e0, e1, 0, a);
    }
}

hipError_t launch_batch_phase(int d, const BatchPhaseArgs& a, int blocks, hipStream_t s, hipEvent_t e0,
                              hipEvent_t e1, char* name, size_t name_len) {
    if (a.narrow)
        launch_batch_phase_st<uint16_t>(d, a, blocks, s, e0, e1, name, name_len);
    else
        launch_batch_phase_st<uint32_t>(d, a, blocks, s, e0, e1, name, name_len);
    return hipGetLastError();
}

static const void* persist_kernel(int d, bool lds2d) {
    if (lds2d) return reinterpret_cast<const void*>(&k_persist_lds2d);
    return d == 2 ? reinterpret_cast<const void*>(&k_persist_global<2>)
                  : reinterpret_cast<const void*>(&k_persist_global<3>);
}

hipError_t persist_max_blocks(int d, bool lds2d, size_t lds_bytes, int* max_blocks) {
    int dev = 0, cus = 0, per_cu = 0;
    hipError_t e;
    if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
    if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess)
        return e;
    if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, persist_kernel(d, lds2d), kBlock,
                                                          lds_bytes)) != hipSuccess)
        return e;
    *max_blocks = per_cu * cus;
    return hipSuccess;
}

hipError_t launch_persist(int d, bool lds2d, const PersistArgs& a, int blocks, size_t lds_bytes,
                          hipStream_t s) {
    PersistArgs copy = a;
    void* args[] = {&copy};
    return hipLaunchCooperativeKernel(persist_kernel(d, lds2d), dim3(blocks), dim3(kBlock), args,
                                      (unsigned int)lds_bytes, s);
}

hipError_t launch_tally(int d, const uint32_t* lat, int lgL, int ncells, int64_t* tab,
                        uint32_t* first, int32_t* anchor, hipStream_t s) {
    const uint32_t nsites = 1u << (d * lgL);
    const int blocks = (int)((nsites + 255) / 256);
    if (d == 2)
        hipLaunchKernelGGL(k_tally_area_perim<2>, dim3(blocks), dim3(256), 0, s, lat, lgL, tab, first);
    else
        hipLaunchKernelGGL(k_tally_area_perim<3>, dim3(blocks), dim3(256), 0, s, lat, lgL, tab, first);
    hipLaunchKernelGGL(k_anchor, dim3((ncells + 1 + 255) / 256), dim3(256), 0, s, first, anchor, ncells, lgL, d);
    return hipGetLastError();
}

hipError_t launch_settle(int64_t* const tab[2], int64_t* const del[2], int G, int ncells,
                         hipStream_t s) {
    hipLaunchKernelGGL(k_settle, dim3((ncells + 1 + 255) / 256), dim3(256), 0, s, tab[(G + 1) & 1],
                       del[(G + 1) & 1], ncells);
    return hipGetLastError();
}

hipError_t launch_csum(int d, const uint32_t* lat, int lgL, const int32_t* anchor, int64_t* csum,
                       hipStream_t s) {
    const uint32_t nsites = 1u << (d * lgL);
    const int blocks = (int)((nsites + 255) / 256);
    if (d == 2)
        hipLaunchKernelGGL(k_tally_csum<2>, dim3(blocks), dim3(256), 0, s, lat, lgL, anchor, csum);
    else
        hipLaunchKernelGGL(k_tally_csum<3>, dim3(blocks), dim3(256), 0, s, lat, lgL, anchor, csum);
    return hipGetLastError();
}

}  // namespace gcpm
