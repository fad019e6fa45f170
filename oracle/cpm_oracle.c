/*
 * cpm_oracle.c -- CPU restatement of the reference Cellular Potts sweep.
 *
 * TEST INFRASTRUCTURE ONLY (see cpm_oracle.h): the checker for the HIP
 * product and the CPU baseline timed beside it.  Plain C99 + optional OpenMP.
 */
#include "cpm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------- RNG -- */

/* random.cu:5-12: x <- 16807 x mod (2^31 - 1), evaluated in double. */
int32_t orc_minstd_next_f64(int32_t seed) {
    const double a = 16807.0;
    const double m = 2147483647.0;
    double temp = (double)seed * a;
    return (int32_t)(temp - m * floor(temp * (1.0 / m)));
}

uint32_t orc_minstd_next_int(uint32_t x) {
    return (uint32_t)(((uint64_t)x * 16807u) % 2147483647u);
}

/* Philox4x32 with 10 rounds (Salmon, Moraes, Dror, Shaw, SC'11, "Parallel
 * random numbers: as easy as 1, 2, 3").  Round: two 32x32->64 products of
 * words 0 and 2 with the multipliers; the key is bumped by the Weyl
 * constants before every round after the first. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t x0 = ctr[0], x1 = ctr[1], x2 = ctr[2], x3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint64_t p0 = (uint64_t)0xD2511F53u * x0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * x2;
        uint32_t y0 = (uint32_t)(p1 >> 32) ^ x1 ^ k0;
        uint32_t y1 = (uint32_t)p1;
        uint32_t y2 = (uint32_t)(p0 >> 32) ^ x3 ^ k1;
        uint32_t y3 = (uint32_t)p0;
        x0 = y0; x1 = y1; x2 = y2; x3 = y3;
    }
    out[0] = x0; out[1] = x1; out[2] = x2; out[3] = x3;
}

/* ------------------------------------------------------- parameters -- */

static int act_on(const orc_params* p) {
    for (int t = 0; t < ORC_NTYPES; t++)
        if (p->lam_act[t] != 0.0) return 1;
    return 0;
}

int64_t orc_act_log2(int v) {
    return v >= 1 ? llround(log2((double)v) * 16777216.0) : 0;
}

int64_t orc_act_multiplier(double lambda, int act_max, int scale) {
    if (act_max < 1) return 0;
    return llround(lambda * (double)scale / (double)act_max * 65536.0);
}

static int64_t act_exp_table(int j) { return llround(exp2((double)j / 256.0) * 1073741824.0); }

int64_t orc_act_geo_mean(int64_t sum_lg, int count) {
    static int64_t ex[256];
    static int64_t ln2q;
    static int init = 0;
    if (!init) {
        for (int j = 0; j < 256; j++) ex[j] = act_exp_table(j);
        ln2q = llround(log(2.0) * 1073741824.0);
        init = 1;
    }
    int64_t mean = sum_lg / count;
    int64_t I = mean >> 24, f = mean & 0xFFFFFF;
    int64_t j = f >> 16, r = f & 0xFFFF;
    int64_t X = (r * ln2q) >> 24;                      /* r ln2 / 2^24 in 2^-30 */
    int64_t P = ((int64_t)1 << 30) + X + ((X * X) >> 31);
    int64_t M = (ex[j] * P) >> 30;                     /* 2^(f / 2^24) in 2^-30 */
    return (M << I) >> 14;
}

static int pers_on(const orc_params* p) {
    for (int t = 0; t < ORC_NTYPES; t++)
        if (p->lam_pers[t] != 0.0) return 1;
    return 0;
}

static int chem_on(const orc_params* p) {
    for (int t = 0; t < ORC_NTYPES; t++)
        if (p->lam_chem[t] != 0.0) return 1;
    return 0;
}

int64_t orc_chem_multiplier(double lambda, int scale, int k) {
    return llround(lambda * (double)scale * (double)(1 << ORC_CHEM_FRAC) / sqrt((double)k));
}

int orc_quantize_field(int64_t n, const double* f, int32_t* out) {
    for (int64_t i = 0; i < n; i++) {
        if (!(fabs(f[i]) < 32768.0)) return -1;
        out[i] = (int32_t)llround(f[i] * (double)(1 << ORC_FIELD_FRAC));
    }
    return 0;
}

int orc_energy_scale(const orc_params* p) {
    /* chemotaxis and act terms are resolved at 1/256 energy units */
    for (int k = chem_on(p) || act_on(p) || pers_on(p) ? ORC_MAX_SCALE_LOG2 : 0;
         k <= ORC_MAX_SCALE_LOG2; k++) {
        double s = (double)(1 << k);
        int ok = 1;
        for (int t = 0; t < ORC_NTYPES && ok; t++) {
            double v[2] = {p->lam_area[t] * s, p->lam_perim[t] * s};
            for (int j = 0; j < 2; j++) {
                if (v[j] != floor(v[j]) || fabs(v[j]) >= 2147483648.0) ok = 0;
            }
        }
        if (ok) return 1 << k;
    }
    return -1;
}

int orc_threshold_table(double temperature, int scale, uint32_t* out, int cap) {
    /* thr[0] is never consulted (dH <= 0 is always accepted). */
    if (cap < 1) return -1;
    out[0] = 0xFFFFFFFFu;
    if (!(temperature > 0.0)) return 1;
    int n = 1;
    for (int k = 1; k < cap; k++) {
        double v = floor(exp(-(double)k / ((double)scale * temperature)) * 4294967296.0);
        if (v <= 0.0) return n;
        out[k] = (uint32_t)v;
        n = k + 1;
    }
    return -2; /* table would exceed cap */
}

/* The reference accepts a positive energy E when r < exp(-E / T) with
 * r = u / (2^31 - 1) and u the next MINSTD value (kernels2d.cu:522-525).  In
 * exact arithmetic that is u < exp(-E/T) (2^31 - 1), i.e. u < ceil(...): with
 * E = k / scale the table holds thr31[k] = ceil(exp(-k / (scale T)) (2^31 -
 * 1)) while it exceeds 1 (u >= 1, so a smaller entry never accepts). */
int orc_ref_threshold_table(double temperature, int scale, uint32_t* out, int cap) {
    if (cap < 1) return -1;
    out[0] = 0xFFFFFFFFu;
    if (!(temperature > 0.0)) return 1;
    int n = 1;
    for (int k = 1; k < cap; k++) {
        double v = ceil(exp(-(double)k / ((double)scale * temperature)) * 2147483647.0);
        if (v <= 1.0) return n;
        out[k] = (uint32_t)v;
        n = k + 1;
    }
    return -2;
}

/* ------------------------------------------------------- geometry -- */

typedef struct {
    int n;
    int off[26][3];
} orc_nbhd;

/* Offsets in lexicographic (dz, dy, dx) order, dx fastest, origin skipped. */
static void make_nbhd(int d, orc_nbhd* nb) {
    nb->n = 0;
    int zlo = d == 3 ? -1 : 0, zhi = d == 3 ? 1 : 0;
    for (int dz = zlo; dz <= zhi; dz++)
        for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
                if (!dx && !dy && !dz) continue;
                nb->off[nb->n][0] = dx;
                nb->off[nb->n][1] = dy;
                nb->off[nb->n][2] = dz;
                nb->n++;
            }
}

static inline int64_t site(int L, int x, int y, int z) {
    return ((int64_t)z * L + y) * L + x;
}

static inline int32_t wrapc(int L, int v) {
    /* signed minimum image in [-L/2, L/2) */
    return ((v + L / 2) & (L - 1)) - L / 2;
}

static int log2i(int v) {
    int k = 0;
    while ((1 << k) < v) k++;
    return k;
}

/* --------------------------------------------------------- tallies -- */

int orc_tally(int L, int d, int ncells, const uint32_t* lat, int32_t* area,
              int32_t* perim, int32_t* anchor, int64_t* csum) {
    if (L < 2 || (L & (L - 1)) || (d != 2 && d != 3)) return -2;
    orc_nbhd nb;
    make_nbhd(d, &nb);
    int Lz = d == 3 ? L : 1;
    memset(area, 0, sizeof(int32_t) * (size_t)(ncells + 1));
    memset(perim, 0, sizeof(int32_t) * (size_t)(ncells + 1));
    memset(csum, 0, sizeof(int64_t) * 3 * (size_t)(ncells + 1));
    int64_t* first = (int64_t*)malloc(sizeof(int64_t) * (size_t)(ncells + 1));
    for (int c = 0; c <= ncells; c++) first[c] = -1;
    for (int z = 0; z < Lz; z++)
        for (int y = 0; y < L; y++)
            for (int x = 0; x < L; x++) {
                uint32_t v = lat[site(L, x, y, z)];
                uint32_t c = v & 0xFFFFFFu;
                if ((v >> 24) >= ORC_NTYPES || c > (uint32_t)ncells) {
                    free(first);
                    return -1;
                }
                if (!c) continue;
                area[c]++;
                if (first[c] < 0) first[c] = site(L, x, y, z);
                int cnt = 0;
                for (int k = 0; k < nb.n; k++) {
                    int nx = (x + nb.off[k][0]) & (L - 1);
                    int ny = (y + nb.off[k][1]) & (L - 1);
                    int nz = d == 3 ? (z + nb.off[k][2]) & (L - 1) : 0;
                    if ((lat[site(L, nx, ny, nz)] & 0xFFFFFFu) != c) cnt++;
                }
                perim[c] += cnt;
            }
    for (int c = 0; c <= ncells; c++) {
        int64_t f = first[c] < 0 ? 0 : first[c];
        anchor[3 * c + 0] = (int32_t)(f % L);
        anchor[3 * c + 1] = (int32_t)((f / L) % L);
        anchor[3 * c + 2] = d == 3 ? (int32_t)(f / ((int64_t)L * L)) : 0;
    }
    for (int z = 0; z < Lz; z++)
        for (int y = 0; y < L; y++)
            for (int x = 0; x < L; x++) {
                uint32_t c = lat[site(L, x, y, z)] & 0xFFFFFFu;
                if (!c) continue;
                csum[3 * c + 0] += wrapc(L, x - anchor[3 * c + 0]);
                csum[3 * c + 1] += wrapc(L, y - anchor[3 * c + 1]);
                if (d == 3) csum[3 * c + 2] += wrapc(L, z - anchor[3 * c + 2]);
            }
    free(first);
    return 0;
}

/* ------------------------------------------------------------ sweep -- */

typedef struct {
    int64_t J[ORC_NTYPES][ORC_NTYPES];   /* J * scale */
    int64_t JS[ORC_NTYPES][ORC_NTYPES];  /* (J + S) * scale */
    int64_t la[ORC_NTYPES], lp[ORC_NTYPES];
    int64_t a0[ORC_NTYPES], p0[ORC_NTYPES];
    int chem;                            /* chemotaxis term on */
    int64_t chem_m[ORC_NTYPES][4];       /* orc_chem_multiplier, k = 1..3 */
    const int32_t* field_q;
    int pers;                            /* persistence term on (3D) */
    int64_t pers_m[ORC_NTYPES][4];
    const int32_t* pref_q;
    int act;                             /* act term on */
    int64_t act_m[ORC_NTYPES];           /* orc_act_multiplier */
    int32_t act_max[ORC_NTYPES];
    uint32_t* act_lat;
} orc_scaled;

static void scale_params(const orc_params* p, int scale, orc_scaled* q) {
    for (int a = 0; a < ORC_NTYPES; a++) {
        for (int b = 0; b < ORC_NTYPES; b++) {
            q->J[a][b] = (int64_t)p->J[a][b] * scale;
            q->JS[a][b] = ((int64_t)p->J[a][b] + p->S[a][b]) * scale;
        }
        q->la[a] = (int64_t)(p->lam_area[a] * scale);
        q->lp[a] = (int64_t)(p->lam_perim[a] * scale);
        q->a0[a] = p->target_area[a];
        q->p0[a] = p->target_perim[a];
        for (int k = 0; k < 4; k++) {
            q->chem_m[a][k] = k ? orc_chem_multiplier(p->lam_chem[a], scale, k) : 0;
            q->pers_m[a][k] = k ? orc_chem_multiplier(p->lam_pers[a], scale, k) : 0;
        }
    }
    q->pers = pers_on(p) && p->pref_q != NULL;
    q->pref_q = p->pref_q;
    q->chem = chem_on(p) && p->field_q != NULL;
    q->field_q = p->field_q;
    q->act = act_on(p) && p->act != NULL;
    q->act_lat = p->act;
    for (int a = 0; a < ORC_NTYPES; a++) {
        q->act_max[a] = p->act_max[a];
        q->act_m[a] = orc_act_multiplier(p->lam_act[a], p->act_max[a], scale);
    }
}

/* Persistence energy (3D) of copying vs into vt through offset o. */
static int64_t pers_energy(const orc_scaled* q, uint32_t vs, uint32_t vt, const int o[3]) {
    uint32_t cs = vs & 0xFFFFFFu, ct = vt & 0xFFFFFFu;
    uint32_t tl = cs ? vs >> 24 : vt >> 24;
    uint32_t cell = cs ? cs : ct;
    int k = (o[0] != 0) + (o[1] != 0) + (o[2] != 0);
    int64_t m = q->pers_m[tl][k];
    if (m == 0 || cell == 0) return 0;
    int64_t g = 0;
    for (int a = 0; a < 3; a++) g += (int64_t)o[a] * q->pref_q[3 * cell + a];
    return (m * g + ((int64_t)1 << 23)) >> 24;
}

/* Geometric mean of act values over site (x, y, z) and its neighbours that
 * hold value v (kernels2d.cu:386-397 / 430-460), in 2^-16 units; 0 for
 * medium or when any value is 0. */
static int64_t act_site_mean(int L, int d, const orc_nbhd* nb, const uint32_t* lat,
                             const orc_scaled* q, int x, int y, int z, uint32_t v,
                             uint32_t tick) {
    if ((v & 0xFFFFFFu) == 0) return 0;
    int64_t am = q->act_max[v >> 24];
    int64_t sum = 0;
    int count = 0, zero = 0;
    for (int k = -1; k < nb->n; k++) {
        int64_t si;
        if (k < 0) {
            si = site(L, x, y, z);
        } else {
            int nx = (x + nb->off[k][0]) & (L - 1);
            int ny = (y + nb->off[k][1]) & (L - 1);
            int nz = d == 3 ? (z + nb->off[k][2]) & (L - 1) : 0;
            si = site(L, nx, ny, nz);
            if (lat[si] != v) continue;
        }
        int64_t a = am + (int64_t)q->act_lat[si] - (int64_t)tick;
        if (a > am) a = am;
        if (a <= 0) zero = 1; else sum += orc_act_log2((int)a);
        count++;
    }
    return zero ? 0 : orc_act_geo_mean(sum, count);
}

/* Act energy of copying vs (at source s) into target t holding vt. */
static int64_t act_energy(int L, int d, const orc_nbhd* nb, const uint32_t* lat,
                          const orc_scaled* q, int tx, int ty, int tz, int sx, int sy, int sz,
                          uint32_t vs, uint32_t vt, uint32_t tick) {
    uint32_t tl = (vt >> 24) ? vt >> 24 : vs >> 24;
    int64_t m = q->act_m[tl];
    if (m == 0) return 0;
    int64_t gs = act_site_mean(L, d, nb, lat, q, sx, sy, sz, vs, tick);
    int64_t gt = act_site_mean(L, d, nb, lat, q, tx, ty, tz, vt, tick);
    return (-m * (gs - gt) + ((int64_t)1 << 31)) >> 32;
}

/* Chemotaxis energy of copying from the neighbour at offset o of target t
 * (see cpm_oracle.h), in units of 1/scale. */
static int64_t chem_energy(int L, int d, const orc_scaled* q, uint32_t ts, uint32_t tt,
                           const int o[3], int64_t src) {
    uint32_t tl = ts ? ts : tt;
    int k = (o[0] != 0) + (o[1] != 0) + (d == 3 && o[2] != 0);
    int64_t m = q->chem_m[tl][k];
    if (m == 0) return 0;
    int64_t n = d == 3 ? (int64_t)L * L * L : (int64_t)L * L;
    int64_t g = 0;
    for (int a = 0; a < d; a++) g += (int64_t)o[a] * q->field_q[a * n + src];
    return (m * g + ((int64_t)1 << 23)) >> 24;
}

/* dH of writing value vs into target t, given its neighbourhood (the energy
 * part of kernels2d.cu:346-519).  as_/at_ return the like-neighbour counts
 * of source and target spins. */
static int64_t delta_h_core(int L, int d, const orc_nbhd* nb, const uint32_t* lat,
                            const int32_t* areaR, const int32_t* perimR,
                            const orc_scaled* q, int tx, int ty, int tz, uint32_t vs,
                            uint32_t vt, int* as_, int* at_) {
    uint32_t ts = vs >> 24, tt = vt >> 24;
    int64_t dh = 0;
    int as = 0, at = 0;
    for (int k = 0; k < nb->n; k++) {
        int nx = (tx + nb->off[k][0]) & (L - 1);
        int ny = (ty + nb->off[k][1]) & (L - 1);
        int nz = d == 3 ? (tz + nb->off[k][2]) & (L - 1) : 0;
        uint32_t vn = lat[site(L, nx, ny, nz)];
        uint32_t tn = vn >> 24;
        if (vn != vs) dh += q->J[ts][tn]; else as++;
        if (vn != vt) dh -= q->JS[tt][tn]; else at++;
    }
    int N = nb->n;
    uint32_t cs = vs & 0xFFFFFFu, ct = vt & 0xFFFFFFu;
    if (cs) {
        int64_t A = areaR[cs], P = perimR[cs];
        int64_t dP = N - 2 * as;
        dh += q->la[ts] * (1 + 2 * A - 2 * q->a0[ts]);
        dh += q->lp[ts] * dP * (dP + 2 * (P - q->p0[ts]));
    }
    if (ct) {
        int64_t A = areaR[ct], P = perimR[ct];
        int64_t dP = 2 * at - N;
        dh += q->la[tt] * (1 - 2 * A + 2 * q->a0[tt]);
        dh += q->lp[tt] * dP * (dP + 2 * (P - q->p0[tt]));
    }
    *as_ = as;
    *at_ = at;
    return dh;
}

int64_t orc_delta_h(int L, int d, const uint32_t* lat, const int32_t* area,
                    const int32_t* perim, const orc_params* p, int scale, int64_t t,
                    int64_t s) {
    orc_nbhd nb;
    make_nbhd(d, &nb);
    orc_scaled q;
    scale_params(p, scale, &q);
    int tx = (int)(t % L), ty = (int)((t / L) % L), tz = d == 3 ? (int)(t / ((int64_t)L * L)) : 0;
    int sx = (int)(s % L), sy = (int)((s / L) % L), sz = d == 3 ? (int)(s / ((int64_t)L * L)) : 0;
    int as, at;
    int64_t dh = delta_h_core(L, d, &nb, lat, area, perim, &q, tx, ty, tz, lat[s], lat[t], &as, &at);
    /* o = s - t, wrapped into {-1, 0, 1} */
    int o[3] = {wrapc(L, sx - tx), wrapc(L, sy - ty), d == 3 ? wrapc(L, sz - tz) : 0};
    if (q.chem) dh += chem_energy(L, d, &q, lat[s] >> 24, lat[t] >> 24, o, s);
    if (q.pers && d == 3) dh += pers_energy(&q, lat[s], lat[t], o);
    return dh;
}

int64_t orc_act_energy(int L, int d, const uint32_t* lat, const orc_params* p, int scale,
                       int64_t t, int64_t s, uint32_t tick) {
    orc_nbhd nb;
    make_nbhd(d, &nb);
    orc_scaled q;
    scale_params(p, scale, &q);
    if (!q.act) return 0;
    int tx = (int)(t % L), ty = (int)((t / L) % L), tz = d == 3 ? (int)(t / ((int64_t)L * L)) : 0;
    int sx = (int)(s % L), sy = (int)((s / L) % L), sz = d == 3 ? (int)(s / ((int64_t)L * L)) : 0;
    return act_energy(L, d, &nb, lat, &q, tx, ty, tz, sx, sy, sz, lat[s], lat[t], tick);
}

static void claim_min(uint32_t* owner, uint32_t c, uint32_t sb) {
    uint32_t cur = __atomic_load_n(&owner[c], __ATOMIC_RELAXED);
    while (sb < cur &&
           !__atomic_compare_exchange_n(&owner[c], &cur, sb, 0, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
    }
}

int orc_sweep(int L, int d, int ncells, uint32_t* lat, int32_t* area, int32_t* perim,
              const int32_t* anchor, int64_t* csum, const orc_params* p, int c, int u,
              uint64_t seed, uint32_t mcs0, int nmcs, int nthreads) {
    return orc_sweep_sync(L, d, ncells, lat, area, perim, anchor, csum, p, c, u, seed, mcs0, nmcs,
                          nthreads, 0);
}

int orc_sweep_sync(int L, int d, int ncells, uint32_t* lat, int32_t* area, int32_t* perim,
                   const int32_t* anchor, int64_t* csum, const orc_params* p, int c, int u,
                   uint64_t seed, uint32_t mcs0, int nmcs, int nthreads, int cell_sync) {
    return orc_sweep_ex(L, d, ncells, lat, area, perim, anchor, csum, p, c, u, seed, mcs0, nmcs,
                        nthreads, cell_sync, 1, 2 * c);
}

int orc_schedule(int L, int d, int c, int shared, int ppt, int* K, int* S) {
    (void)d;
    if (c < 1 || (c & (c - 1))) return -3;
    int R = shared ? 2 * c : ppt;
    if (R < c || R % c || L % R) return -3;
    *K = R / c;
    int s = 1;
    if (!shared) {
        /* same-group sub-blocks must be >= 3 sites apart: a non-shared
         * attempt writes up to one site outside its sub-block and reads up
         * to two */
        while (s * R - c < 3) s *= 2;
        if (s > L / R) s = L / R;
    }
    *S = s;
    return 0;
}

/* Reference stream (random.cu:5-12, kernels2d.cu:99,199-243,522,
 * kernels3d.cu:86,160-215): one MINSTD state per region of R sites per axis
 * (a reference thread), seeded at the start of every call (one reference
 * launch) with x = MINSTD(index of the region's first site + seed32) and then
 * consumed in the region's attempt order: the position in the sub-block (d
 * draws mod c), then direction triples (each axis u mod 3 - 1) until one is
 * not rejected, then -- when the two sites differ and the target type is not
 * fixed -- the acceptance draw u.  The rejection test is the reference's:
 * all drawn coordinates equal in 2D and in 3D with shared roles; in 3D
 * without shared roles the reference compares the z coordinate with y
 * (kernels3d.cu:208), which is kept.  At most kRefTries direction triples
 * are drawn (a stream stuck at one value would otherwise never end).
 * Acceptance: dH <= 0 or u < thr31[dH] (orc_ref_threshold_table). */
#define REF_TRIES 256
static int32_t ref_next(int32_t x) { return orc_minstd_next_f64(x); }

static int sweep_impl(int L, int d, int ncells, uint32_t* lat, int32_t* area, int32_t* perim,
                      const int32_t* anchor, int64_t* csum, const orc_params* p, int c, int u,
                      uint64_t seed, uint32_t mcs0, int nmcs, int nthreads, int cell_sync,
                      int shared, int ppt, int refstream, int64_t* trace, int64_t trace_cap,
                      int64_t* trace_n, int live, int64_t* counts, int tpb);

int orc_sweep_ex(int L, int d, int ncells, uint32_t* lat, int32_t* area, int32_t* perim,
                 const int32_t* anchor, int64_t* csum, const orc_params* p, int c, int u,
                 uint64_t seed, uint32_t mcs0, int nmcs, int nthreads, int cell_sync, int shared,
                 int ppt) {
    return sweep_impl(L, d, ncells, lat, area, perim, anchor, csum, p, c, u, seed, mcs0, nmcs,
                      nthreads, cell_sync, shared, ppt, 0, NULL, 0, NULL, 0, NULL, 1);
}

/* Live-table checkerboard: the same schedule, stream and decisions as
 * orc_sweep_ex, except that each phase's attempts are made one at a time in
 * a per-phase pseudo-random order (qi -> (a qi + b) mod nq, a odd, from the
 * phase's own Philox draw), every attempt reading the area and perimeter
 * tables as the earlier attempts of the phase left them.  It is the other
 * end of what the reference's concurrent phase can produce (kernels2d.cu:
 * 292-302 read cellArea / cellPerimeter while other threads' atomicAdds at
 * :552-569 land): orc_sweep reads every table entry before any write of the
 * phase, this reads every write that precedes it in one sequential order.
 * Single-threaded; used only for statistical comparisons. */
int orc_sweep_live(int L, int d, int ncells, uint32_t* lat, int32_t* area, int32_t* perim,
                   const int32_t* anchor, int64_t* csum, const orc_params* p, int c, int u,
                   uint64_t seed, uint32_t mcs0, int nmcs, int shared, int ppt) {
    return sweep_impl(L, d, ncells, lat, area, perim, anchor, csum, p, c, u, seed, mcs0, nmcs, 1, 0,
                      shared, ppt, 0, NULL, 0, NULL, 1, NULL, 1);
}

/* orc_sweep_ex that also counts, once per attempt (in the commit / apply pass
 * of a claimed or split phase), into counts[4]: attempts whose source and
 * target differ, attempts evaluated after the fixed-type rules, proposals
 * accepted, flips committed (the reference's analytics block, whose
 * increments at kernels2d.cu:292,526,531 are commented out). */
int orc_sweep_counts(int L, int d, int ncells, uint32_t* lat, int32_t* area, int32_t* perim,
                     const int32_t* anchor, int64_t* csum, const orc_params* p, int c, int u,
                     uint64_t seed, uint32_t mcs0, int nmcs, int nthreads, int cell_sync, int shared,
                     int ppt, int64_t* counts) {
    return sweep_impl(L, d, ncells, lat, area, perim, anchor, csum, p, c, u, seed, mcs0, nmcs, nthreads,
                      cell_sync, shared, ppt, 0, NULL, 0, NULL, 0, counts, 1);
}

int orc_sweep_refstream(int L, int d, int ncells, uint32_t* lat, int32_t* area, int32_t* perim,
                        const int32_t* anchor, int64_t* csum, const orc_params* p, int c, int u,
                        uint64_t seed, uint32_t mcs0, int nmcs, int nthreads, int cell_sync,
                        int shared, int ppt, int64_t* trace, int64_t trace_cap, int64_t* trace_n,
                        int tpb) {
    /* the reference's shared mode with ppt > 2c: global blocks of tpb threads
     * per axis, each thread looping over ppt / 2c regions (kernels2d.cu:72-86,
     * 92-106, cpm.cu:363) */
    if (shared && ppt != 2 * c &&
        (ppt < 2 * c || ppt % (2 * c) || tpb < 1 || L % ppt || (L / ppt) % tpb))
        return -9;
    return sweep_impl(L, d, ncells, lat, area, perim, anchor, csum, p, c, u, seed, mcs0, nmcs,
                      trace ? 1 : nthreads, cell_sync, shared, ppt, 1, trace, trace_cap, trace_n, 0, NULL,
                      tpb);
}

static int sweep_impl(int L, int d, int ncells, uint32_t* lat, int32_t* area, int32_t* perim,
                      const int32_t* anchor, int64_t* csum, const orc_params* p, int c, int u,
                      uint64_t seed, uint32_t mcs0, int nmcs, int nthreads, int cell_sync,
                      int shared, int ppt, int refstream, int64_t* trace, int64_t trace_cap,
                      int64_t* trace_n, int live, int64_t* counts, int tpb) {
    if (d != 2 && d != 3) return -1;
    /* Act with shared roles and c = 1: an attempt's act term reads the
     * source's neighbourhood, two sites from its target, where other targets
     * of the phase lie.  Each phase is then made in two passes: every
     * attempt is decided against the phase-start lattice (nothing written),
     * then the accepted ones are applied (a target's own neighbourhood holds
     * no other target, so the applied deltas are those decided).  With
     * cell_sync the decide pass also claims and the apply pass checks the
     * claims (kernels.cu:123-141 as a deterministic rule). */
    const int split = shared && act_on(p) && p->act != NULL && c < 2;
    if (live && (split || cell_sync || refstream)) return -10; /* (not needed for the comparisons) */
    if (live) nthreads = 1;
    if (L < 2 || (L & (L - 1))) return -2;
    int K, S;
    if (orc_schedule(L, d, c, shared, ppt, &K, &S)) return -3;
    const int R = K * c;
    int cd = d == 3 ? c * c * c : c * c;
    if (u < 1 || u > cd) return -4;
    int scale = orc_energy_scale(p);
    if (scale < 0) return -5;
    int cap = 1 << 22;
    uint32_t* thr = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)cap);
    int nthr = refstream ? orc_ref_threshold_table(p->temperature, scale, thr, cap)
                         : orc_threshold_table(p->temperature, scale, thr, cap);
    if (nthr < 0) {
        free(thr);
        return -6;
    }
    orc_scaled q;
    scale_params(p, scale, &q);
    orc_nbhd nb;
    make_nbhd(d, &nb);
    const int N = nb.n;
    const int lc = log2i(c);
    /* The reference's shared mode with ppt > 2c (reference stream): the
     * lattice is cut into blocks of tpb threads per axis, each thread owning
     * G = ppt / 2c regions of 2c sites per axis; a block's threads work on one
     * slice of G^d at a time (blockx, blocky(, blockz), x outermost), so thread
     * j of the colour phase in slice g works in region
     * (j / tpb) G tpb + g tpb + j mod tpb (kernels2d.cu:72-86,92-106;
     * kernels3d.cu:65-108).  One state per thread, seeded from its first
     * site j ppt (kernels2d.cu:70-71,99). */
    const int gb = refstream && shared && ppt > R;
    const int G = gb ? ppt / R : 1;
    const int tb = gb ? tpb : 1;
    const int T = gb ? L / ppt : 0;  /* threads per axis */
    const int ngb = d == 3 ? G * G * G : G * G;
    /* reference stream: one state per thread (a region, or with gb a thread
     * of ppt sites per axis), seeded per call */
    const int nreg = gb ? T : L / R;
    const int sreg = gb ? ppt : R;
    int32_t* rstate = NULL;
    if (refstream) {
        int64_t nr = d == 3 ? (int64_t)nreg * nreg * nreg : (int64_t)nreg * nreg;
        rstate = (int32_t*)malloc(sizeof(int32_t) * (size_t)nr);
        for (int64_t ri = 0; ri < nr; ri++) {
            int rx = (int)(ri % nreg), ry = (int)((ri / nreg) % nreg);
            int rz = d == 3 ? (int)(ri / ((int64_t)nreg * nreg)) : 0;
            uint32_t base = (uint32_t)site(L, rx * sreg, ry * sreg, rz * sreg); /* kernels2d.cu:99 */
            rstate[ri] = ref_next((int32_t)(base + (uint32_t)seed));
        }
        if (trace_n) *trace_n = 0;
    }
    const uint32_t cm = (uint32_t)c - 1;
    const int nsb = L / c;       /* sub-blocks per axis */
    const int h = gb ? T : L / R / S;  /* same-phase sub-blocks per axis */
    const int64_t nq = d == 3 ? (int64_t)h * h * h : (int64_t)h * h;
    const int rounds = cd / u;   /* kernels2d.cu:95 */
    const int ncol = d == 3 ? K * K * K : K * K;   /* colours: K^d (kernels2d.cu:91) */
    const int ngrp = d == 3 ? S * S * S : S * S;
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};

    int32_t* areaR = (int32_t*)malloc(sizeof(int32_t) * (size_t)(ncells + 1));
    int32_t* perimR = (int32_t*)malloc(sizeof(int32_t) * (size_t)(ncells + 1));
    uint32_t* owner = cell_sync ? (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(ncells + 1)) : NULL;
    uint8_t* dec = split ? (uint8_t*)malloc((size_t)nq) : NULL;
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#else
    (void)nthreads;
#endif

    for (int m = 0; m < nmcs; m++) {
        uint32_t mcs = mcs0 + (uint32_t)m;
        for (int r = 0; r < rounds; r++) {
            for (int gs = 0; gs < ngb; gs++) {  /* the global-block slice (one unless gb) */
            const int gsx = d == 3 ? gs / (G * G) : gs / G;
            const int gsy = d == 3 ? (gs / G) % G : gs % G;
            const int gsz = d == 3 ? gs % G : 0;
            for (int pc = 0; pc < ncol * ngrp; pc++) {
                /* colour (cbX outermost, kernels2d.cu:179-180), then group */
                int col = pc / ngrp, grp = pc % ngrp;
                int cx = d == 3 ? col / (K * K) : col / K;
                int cy = d == 3 ? (col / K) % K : col % K;
                int cz = d == 3 ? col % K : 0;
                int gx = d == 3 ? grp / (S * S) : grp / S;
                int gy = d == 3 ? (grp / S) % S : grp % S;
                int gz = d == 3 ? grp % S : 0;
                uint32_t phase = (uint32_t)((r * ngb + gs) * ncol * ngrp + pc);
                /* the phase-start snapshot the attempts read (live: the tables themselves) */
                const int32_t* aR = live ? area : areaR;
                const int32_t* pR = live ? perim : perimR;
                uint64_t perm_a = 1, perm_b = 0;
                if (live) {
                    uint32_t pw[4];
                    const uint32_t pctr[4] = {0xA11CE5u, mcs, phase, 0x11FEu};
                    orc_philox4x32_10(pctr, key, pw);
                    perm_a = ((uint64_t)pw[0] << 1) | 1u;
                    perm_b = pw[1];
                } else {
                    memcpy(areaR, area, sizeof(int32_t) * (size_t)(ncells + 1));
                    memcpy(perimR, perim, sizeof(int32_t) * (size_t)(ncells + 1));
                }
                /* cell_sync (kernels.cu:123-141, as a deterministic rule), per
                 * attempt index i: pass 0 claims the cells of every accepted
                 * proposal i for the lowest sub-block index; pass 1 recomputes
                 * each attempt i (nothing has been written) and commits it only
                 * if its sub-block holds the claim on both of its non-medium
                 * cells.  Without cell_sync one pass makes all u attempts. */
                for (int i0 = 0; i0 < u; i0 += cell_sync ? 1 : u)
                for (int pass = cell_sync || split ? 0 : 1; pass < 2; pass++) {
                if (cell_sync && pass == 0) memset(owner, 0xFF, sizeof(uint32_t) * (size_t)(ncells + 1));
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads) schedule(static)
#endif
                for (int64_t qj = 0; qj < nq; qj++) {
                    /* (nq is a power of two: an odd multiplier permutes) */
                    const int64_t qi = live ? (int64_t)((perm_a * (uint64_t)qj + perm_b) & (uint64_t)(nq - 1)) : qj;
                    /* sub-block of region (a S + g) at colour offset c */
                    int bx = ((int)(qi % h) * S + gx) * K + cx;
                    int by = ((int)((qi / h) % h) * S + gy) * K + cy;
                    int bz = d == 3 ? ((int)(qi / ((int64_t)h * h)) * S + gz) * K + cz : 0;
                    int64_t rid = d == 3 ? ((int64_t)(bz / K) * nreg + by / K) * nreg + bx / K
                                         : (int64_t)(by / K) * nreg + bx / K;
                    if (gb) {  /* thread j of slice gs: its region, and its own state */
                        const int jx = (int)(qi % h), jy = (int)((qi / h) % h);
                        const int jz = d == 3 ? (int)(qi / ((int64_t)h * h)) : 0;
                        bx = ((jx / tb) * G * tb + gsx * tb + jx % tb) * K + cx;
                        by = ((jy / tb) * G * tb + gsy * tb + jy % tb) * K + cy;
                        bz = d == 3 ? ((jz / tb) * G * tb + gsz * tb + jz % tb) * K + cz : 0;
                        rid = qi;
                    }
                    uint32_t sb = (uint32_t)(((int64_t)bz * nsb + by) * nsb + bx);
                    int32_t rs = refstream ? rstate[rid] : 0;
                    if (split && pass == 0) dec[qi] = 0;
                    for (int i = i0; i < (cell_sync ? i0 + 1 : u); i++) {
                        uint32_t w[4] = {0, 0, 0, 0};
                        int px, py, pz, qx, qy, qz;
                        int o[3]; /* the source's offset from the target */
                        if (!refstream) {
                            /* word 0: position, 1: neighbour, 2: acceptance.
                             * With c = 1 only words 1 and 2 are needed: the
                             * same-colour sub-blocks 2j and 2j+1 along x share
                             * the even one's block, the odd one taking words 3
                             * and 0 (the build's own stream layout; the
                             * reference draws MINSTD per thread) */
                            const int odd = lc == 0 ? (int)((qi % h) & 1) : 0;
                            uint32_t ctr[4] = {sb - (uint32_t)(odd * S * K), mcs, phase, (uint32_t)i};
                            orc_philox4x32_10(ctr, key, w);
                            if (odd) {
                                w[1] = w[3];
                                w[2] = w[0];
                            }
                            /* the position drawn in the sub-block and a
                             * neighbour: shared roles (kernels2d.cu:204-226):
                             * target, source; non-shared (:227-243): source,
                             * target */
                            px = bx * c + (int)(w[0] & cm);
                            py = by * c + (int)((w[0] >> lc) & cm);
                            pz = d == 3 ? bz * c + (int)((w[0] >> (2 * lc)) & cm) : 0;
                            int dn = (int)(((uint64_t)w[1] * (uint64_t)N) >> 32);
                            qx = (px + nb.off[dn][0]) & (L - 1);
                            qy = (py + nb.off[dn][1]) & (L - 1);
                            qz = d == 3 ? (pz + nb.off[dn][2]) & (L - 1) : 0;
                            /* (the list is symmetric: entry N-1-n is minus entry n) */
                            int dir = shared ? dn : N - 1 - dn;
                            for (int a = 0; a < 3; a++) o[a] = nb.off[dir][a];
                        } else {
                            rs = ref_next(rs);
                            px = bx * c + rs % c;
                            rs = ref_next(rs);
                            py = by * c + rs % c;
                            pz = 0;
                            if (d == 3) {
                                rs = ref_next(rs);
                                pz = bz * c + rs % c;
                            }
                            int dx = 0, dy = 0, dz = 0;
                            for (int tries = 0; tries < REF_TRIES; tries++) {
                                rs = ref_next(rs);
                                dx = rs % 3 - 1;
                                rs = ref_next(rs);
                                dy = rs % 3 - 1;
                                if (d == 3) {
                                    rs = ref_next(rs);
                                    dz = rs % 3 - 1;
                                }
                                qx = (px + dx) & (L - 1);
                                qy = (py + dy) & (L - 1);
                                qz = d == 3 ? (pz + dz) & (L - 1) : 0;
                                int rej = d == 2   ? (qx == px && qy == py)
                                          : shared ? (qx == px && qy == py && qz == pz)
                                                   : (qx == px && qy == py && qz == py);
                                if (!rej) break;
                            }
                            o[0] = shared ? dx : -dx;
                            o[1] = shared ? dy : -dy;
                            o[2] = shared ? dz : -dz;
                        }
                        int tx = shared ? px : qx, ty = shared ? py : qy, tz = shared ? pz : qz;
                        int sx = shared ? qx : px, sy = shared ? qy : py, sz = shared ? qz : pz;
                        int64_t ti = site(L, tx, ty, tz);
                        uint32_t vs = lat[site(L, sx, sy, sz)];
                        uint32_t vt = lat[ti];
                        if (refstream) {
                            /* the acceptance draw is taken when the sites
                             * differ and the target is not fixed
                             * (kernels2d.cu:276-284,522) */
                            int take = vs != vt && !p->fixed[vt >> 24];
                            if (take) {
                                rs = ref_next(rs);
                                w[2] = (uint32_t)rs;
                            }
                            if (trace && !((cell_sync || split) && pass == 0) && *trace_n < trace_cap) {
                                int64_t* e = trace + 4 * *trace_n;
                                e[0] = rid;
                                e[1] = ti;
                                e[2] = site(L, sx, sy, sz);
                                e[3] = take ? (int64_t)(uint32_t)rs : -1;
                                ++*trace_n;
                            }
                        }
                        if (vs == vt) continue; /* kernels2d.cu:276 */
                        const int counting = counts != NULL && pass == 1;
                        if (counting) {
#ifdef _OPENMP
#pragma omp atomic
#endif
                            counts[0]++;
                        }
                        /* fixed types (kernels2d.cu:281-290): a fixed target is
                         * never overwritten; a fixed source copies medium.
                         * Onto a medium target that is a medium -> medium
                         * attempt, which the reference evaluates, draws and,
                         * if accepted, stamps act[target] (:532-534); it
                         * changes nothing else, so without the act term it
                         * is skipped */
                        if (p->fixed[vt >> 24]) continue;
                        if (p->fixed[vs >> 24]) {
                            vs = 0;
                            if (vs == vt && !q.act) continue;
                        }
                        if (counting) {
#ifdef _OPENMP
#pragma omp atomic
#endif
                            counts[1]++;
                        }
                        int as, at;
                        int64_t dh = delta_h_core(L, d, &nb, lat, aR, pR, &q, tx, ty,
                                                  tz, vs, vt, &as, &at);
                        int accept;
                        if (split && pass == 1) {
                            accept = dec[qi]; /* decided against the phase-start lattice */
                        } else {
                            if (q.chem)
                                dh += chem_energy(L, d, &q, vs >> 24, vt >> 24, o, site(L, sx, sy, sz));
                            if (q.act)
                                dh += act_energy(L, d, &nb, lat, &q, tx, ty, tz, sx, sy, sz, vs, vt, mcs);
                            if (q.pers && d == 3) dh += pers_energy(&q, vs, vt, o);
                            accept = dh <= 0 || (dh < nthr && w[2] < thr[dh]);
                        }
                        if (split && pass == 0) {
                            dec[qi] = (uint8_t)accept;
                            /* with cell_sync the accepted ones also claim */
                            if (!cell_sync) continue;
                        }
                        if (!accept) continue;
                        if (counting) {
#ifdef _OPENMP
#pragma omp atomic
#endif
                            counts[2]++;
                        }
                        uint32_t cs = vs & 0xFFFFFFu, ct = vt & 0xFFFFFFu;
                        if (cell_sync && pass == 0) {
                            if (cs) claim_min(owner, cs, sb);
                            if (ct) claim_min(owner, ct, sb);
                            continue;
                        }
                        if (cell_sync && ((cs && owner[cs] != sb) || (ct && owner[ct] != sb))) continue;
                        if (counting) {
#ifdef _OPENMP
#pragma omp atomic
#endif
                            counts[3]++;
                        }
                        lat[ti] = vs; /* kernels2d.cu:548 */
                        if (q.act) q.act_lat[ti] = mcs; /* kernels2d.cu:532-534 */
                        int pos[3] = {tx, ty, tz};
                        if (cs) { /* kernels2d.cu:552-560 */
#ifdef _OPENMP
#pragma omp atomic
#endif
                            area[cs] += 1;
#ifdef _OPENMP
#pragma omp atomic
#endif
                            perim[cs] += N - 2 * as;
                            for (int a = 0; a < d; a++) {
                                int64_t w_ = wrapc(L, pos[a] - anchor[3 * cs + a]);
#ifdef _OPENMP
#pragma omp atomic
#endif
                                csum[3 * cs + a] += w_;
                            }
                        }
                        if (ct) { /* kernels2d.cu:561-569 */
#ifdef _OPENMP
#pragma omp atomic
#endif
                            area[ct] -= 1;
#ifdef _OPENMP
#pragma omp atomic
#endif
                            perim[ct] += 2 * at - N;
                            for (int a = 0; a < d; a++) {
                                int64_t w_ = wrapc(L, pos[a] - anchor[3 * ct + a]);
#ifdef _OPENMP
#pragma omp atomic
#endif
                                csum[3 * ct + a] -= w_;
                            }
                        }
                    }
                    /* a claim pass recomputes the attempts its commit pass makes */
                    if (refstream && !((cell_sync || split) && pass == 0)) rstate[rid] = rs;
                }
                }
            }
            }  /* gs */
        }
    }
    free(areaR);
    free(perimR);
    free(owner);
    free(dec);
    free(thr);
    free(rstate);
    return 0;
}

/* Serial random-sequential CPM (the dynamics the reference's parallel
 * sweep approximates): each MCS makes L^d attempts, each at a uniformly
 * drawn target with a uniformly drawn neighbour as source, evaluated and
 * committed one at a time against the live tables.  Same Hamiltonian and
 * acceptance table as orc_sweep; its own stream (Philox keyed by seed and
 * the attempt counter).  Used only for statistical comparisons. */
int orc_serial_sweep(int L, int d, int ncells, uint32_t* lat, int32_t* area, int32_t* perim,
                     const int32_t* anchor, int64_t* csum, const orc_params* p, uint64_t seed,
                     uint32_t mcs0, int nmcs) {
    if (d != 2 && d != 3) return -1;
    if (L < 2 || (L & (L - 1))) return -2;
    int scale = orc_energy_scale(p);
    if (scale < 0) return -5;
    int cap = 1 << 22;
    uint32_t* thr = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)cap);
    int nthr = orc_threshold_table(p->temperature, scale, thr, cap);
    if (nthr < 0) {
        free(thr);
        return -6;
    }
    orc_scaled q;
    scale_params(p, scale, &q);
    orc_nbhd nb;
    make_nbhd(d, &nb);
    const int N = nb.n;
    const int lgL = log2i(L);
    const uint32_t key[2] = {(uint32_t)seed ^ 0x5E71A1u, (uint32_t)(seed >> 32)};
    const int64_t sites = d == 3 ? (int64_t)L * L * L : (int64_t)L * L;
    for (int m = 0; m < nmcs; m++) {
        uint32_t mcs = mcs0 + (uint32_t)m;
        for (int64_t a = 0; a < sites; a++) {
            uint32_t ctr[4] = {(uint32_t)a, (uint32_t)(a >> 32), mcs, 0x5E71A1u};
            uint32_t w[4];
            orc_philox4x32_10(ctr, key, w);
            int tx = (int)(w[0] & (uint32_t)(L - 1));
            int ty = (int)((w[0] >> lgL) & (uint32_t)(L - 1));
            int tz = d == 3 ? (int)(w[3] & (uint32_t)(L - 1)) : 0;
            int dir = (int)(((uint64_t)w[1] * (uint64_t)N) >> 32);
            int sx = (tx + nb.off[dir][0]) & (L - 1);
            int sy = (ty + nb.off[dir][1]) & (L - 1);
            int sz = d == 3 ? (tz + nb.off[dir][2]) & (L - 1) : 0;
            int64_t ti = site(L, tx, ty, tz);
            uint32_t vs = lat[site(L, sx, sy, sz)];
            uint32_t vt = lat[ti];
            if (vs == vt) continue;
            if (p->fixed[vt >> 24]) continue;
            if (p->fixed[vs >> 24]) {
                vs = 0; /* medium -> medium onto a medium target: see orc_sweep */
                if (vs == vt && !q.act) continue;
            }
            int as, at;
            int64_t dh = delta_h_core(L, d, &nb, lat, area, perim, &q, tx, ty, tz, vs, vt, &as, &at);
            if (q.chem)
                dh += chem_energy(L, d, &q, vs >> 24, vt >> 24, nb.off[dir], site(L, sx, sy, sz));
            if (q.act) dh += act_energy(L, d, &nb, lat, &q, tx, ty, tz, sx, sy, sz, vs, vt, mcs);
            if (q.pers && d == 3) dh += pers_energy(&q, vs, vt, nb.off[dir]);
            if (!(dh <= 0 || (dh < nthr && w[2] < thr[dh]))) continue;
            lat[ti] = vs;
            if (q.act) q.act_lat[ti] = mcs;
            uint32_t cs = vs & 0xFFFFFFu, ct = vt & 0xFFFFFFu;
            int pos[3] = {tx, ty, tz};
            if (cs) {
                area[cs] += 1;
                perim[cs] += N - 2 * as;
                for (int k = 0; k < d; k++) csum[3 * cs + k] += wrapc(L, pos[k] - anchor[3 * cs + k]);
            }
            if (ct) {
                area[ct] -= 1;
                perim[ct] += 2 * at - N;
                for (int k = 0; k < d; k++) csum[3 * ct + k] -= wrapc(L, pos[k] - anchor[3 * ct + k]);
            }
        }
    }
    free(thr);
    return 0;
}

/* ----------------------------------------------------------- energy -- */

int orc_energy_stats(int L, int d, int ncells, const uint32_t* lat, const orc_params* p,
                     int64_t* pairs, int64_t* sa2, int64_t* sa1, int64_t* sp2,
                     int64_t* sp1) {
    if (L < 2 || (L & (L - 1)) || (d != 2 && d != 3)) return -2;
    orc_nbhd nb;
    make_nbhd(d, &nb);
    int Lz = d == 3 ? L : 1;
    memset(pairs, 0, sizeof(int64_t) * ORC_NTYPES * ORC_NTYPES);
    memset(sa2, 0, sizeof(int64_t) * ORC_NTYPES);
    memset(sa1, 0, sizeof(int64_t) * ORC_NTYPES);
    memset(sp2, 0, sizeof(int64_t) * ORC_NTYPES);
    memset(sp1, 0, sizeof(int64_t) * ORC_NTYPES);
    /* every unordered pair is met twice over all (site, offset) pairs */
    int64_t* twice = (int64_t*)calloc(ORC_NTYPES * ORC_NTYPES, sizeof(int64_t));
    int32_t* area = (int32_t*)calloc((size_t)ncells + 1, sizeof(int32_t));
    int32_t* perim = (int32_t*)calloc((size_t)ncells + 1, sizeof(int32_t));
    int32_t* ctype = (int32_t*)calloc((size_t)ncells + 1, sizeof(int32_t));
    for (int z = 0; z < Lz; z++)
        for (int y = 0; y < L; y++)
            for (int x = 0; x < L; x++) {
                uint32_t v = lat[site(L, x, y, z)];
                uint32_t c = v & 0xFFFFFFu, t = v >> 24;
                if (c > (uint32_t)ncells || t >= ORC_NTYPES) {
                    free(twice); free(area); free(perim); free(ctype);
                    return -1;
                }
                if (c) { area[c]++; ctype[c] = (int32_t)t; }
                for (int k = 0; k < nb.n; k++) {
                    int nx = (x + nb.off[k][0]) & (L - 1);
                    int ny = (y + nb.off[k][1]) & (L - 1);
                    int nz = d == 3 ? (z + nb.off[k][2]) & (L - 1) : 0;
                    uint32_t vn = lat[site(L, nx, ny, nz)];
                    if (vn == v) continue;
                    uint32_t tn = vn >> 24;
                    uint32_t a = t < tn ? t : tn, b = t < tn ? tn : t;
                    twice[a * ORC_NTYPES + b]++;
                    if (c && (vn & 0xFFFFFFu) != c) perim[c]++;
                }
            }
    for (int i = 0; i < ORC_NTYPES * ORC_NTYPES; i++) pairs[i] = twice[i] / 2;
    for (int c = 1; c <= ncells; c++) {
        if (!area[c]) continue;
        int t = ctype[c];
        int64_t da = (int64_t)area[c] - p->target_area[t];
        int64_t dp = (int64_t)perim[c] - p->target_perim[t];
        sa2[t] += da * da;
        sa1[t] += da;
        sp2[t] += dp * dp;
        sp1[t] += dp;
    }
    free(twice); free(area); free(perim); free(ctype);
    return 0;
}

double orc_hamiltonian(int L, int d, int ncells, const uint32_t* lat, const orc_params* p) {
    int64_t pairs[ORC_NTYPES * ORC_NTYPES], sa2[ORC_NTYPES], sa1[ORC_NTYPES],
        sp2[ORC_NTYPES], sp1[ORC_NTYPES];
    if (orc_energy_stats(L, d, ncells, lat, p, pairs, sa2, sa1, sp2, sp1)) return NAN;
    double H = 0.0;
    for (int a = 0; a < ORC_NTYPES; a++)
        for (int b = a; b < ORC_NTYPES; b++)
            H += (double)p->J[a][b] * (double)pairs[a * ORC_NTYPES + b];
    for (int t = 0; t < ORC_NTYPES; t++)
        H += p->lam_area[t] * (double)sa2[t] + p->lam_perim[t] * (double)sp2[t];
    return H;
}

/* -------------------------------------------------------- readback -- */

void orc_centroids(int L, int d, int ncells, const int32_t* area, const int32_t* anchor,
                   const int64_t* csum, double* out) {
    for (int c = 0; c <= ncells; c++)
        for (int a = 0; a < d; a++) {
            if (!area[c]) {
                out[c * d + a] = NAN;
                continue;
            }
            double v = (double)anchor[3 * c + a] + (double)csum[3 * c + a] / (double)area[c];
            if (v < 0.0) v += L;
            if (v >= L) v -= L;
            out[c * d + a] = v;
        }
}

/* ------------------------------------------------------ persistence -- */

void orc_pref_centroids(int L, int d, int ncells, const int32_t* area, const int32_t* anchor,
                        const int64_t* csum, double* out) {
    (void)L;
    for (int c = 0; c <= ncells; c++)
        for (int a = 0; a < 3; a++)
            out[3 * c + a] = (a < d && area[c] > 0)
                                 ? (double)anchor[3 * c + a] + (double)csum[3 * c + a] / (double)area[c]
                                 : 0.0;
}

static double wrap_half(double x, int L) {
    /* kernels.cu:150-156 */
    if (x > L / 2) x -= L;
    if (x < -L / 2) x += L;
    return x;
}

void orc_update_pref_dirs(int L, int d, int ncells, const int32_t* area, const int32_t* anchor,
                          const int64_t* csum, const int32_t* ctype, const double* diffusion,
                          double* pref, double* ring, int H, int* active, int32_t* pref_q) {
    const size_t rows = (size_t)ncells + 1;
    double* cur = ring + (size_t)(*active) * rows * 3;
    orc_pref_centroids(L, d, ncells, area, anchor, csum, cur);
    double* now = (double*)malloc(sizeof(double) * rows * 3);
    memcpy(now, cur, sizeof(double) * rows * 3);
    *active = (*active + 1) % H;
    const double* prev = ring + (size_t)(*active) * rows * 3;
    for (int c = 1; c <= ncells; c++) {
        if (area[c] <= 0) continue;
        double u[3];
        for (int a = 0; a < 3; a++) u[a] = wrap_half(now[3 * c + a] - prev[3 * c + a], L);
        double len = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
        if (len == 0.0) continue;
        double D = diffusion[ctype[c]];
        double np[3];
        for (int a = 0; a < 3; a++) np[a] = u[a] / len * (1.0 - D) + pref[3 * c + a] * D;
        double nl = sqrt(np[0] * np[0] + np[1] * np[1] + np[2] * np[2]);
        if (nl == 0.0) continue;
        for (int a = 0; a < 3; a++) pref[3 * c + a] = np[a] / nl;
    }
    free(now);
    for (size_t i = 0; i < rows * 3; i++) pref_q[i] = (int32_t)llround(pref[i] * 65536.0);
}

void orc_init_pref_dirs(int ncells, uint64_t seed, double* pref) {
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (int c = 0; c <= ncells; c++) {
        for (uint32_t j = 0;; j++) {
            uint32_t ctr[4] = {(uint32_t)c, 0x50524546u, j, 0u};
            uint32_t w[4];
            orc_philox4x32_10(ctr, key, w);
            double x = (double)w[0] / 2147483648.0 - 1.0;
            double y = (double)w[1] / 2147483648.0 - 1.0;
            double z = (double)w[2] / 2147483648.0 - 1.0;
            double s2 = x * x + y * y + z * z;
            if (s2 > 0.0 && s2 <= 1.0) {
                double n = sqrt(s2);
                pref[3 * c + 0] = x / n;
                pref[3 * c + 1] = y / n;
                pref[3 * c + 2] = z / n;
                break;
            }
        }
    }
}
