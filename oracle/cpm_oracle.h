/*
 * cpm_oracle.h -- CPU restatement of the reference Cellular Potts sweep.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this code, and only as the checker
 * (or as the timed CPU baseline).  The product library never links it.
 *
 * Semantics follow the reference (shabaz/gpu-cpm, src/):
 *   - lattice encoding  id = cell | type << 24           kernels.cu:44-46, cpm.cu:163
 *   - row-major index, x fastest                          kernels.cu:20-42
 *   - Moore 8 / 26 periodic neighbourhood                 kernels2d.cu:359-415, kernels3d.cu:368-429
 *   - adhesion + stickiness, area, perimeter dH           kernels2d.cu:399-412,477-519
 *   - tally recount (area, perimeter)                     kernels2d.cu:5-41, kernels3d.cu:3-47
 *   - shared-mode roles: target drawn inside the
 *     checkerboard sub-block, source a random neighbour  kernels2d.cu:204-226
 *   - colour order cbX outer, cbY (, cbZ) inner,
 *     (c^d / u) rounds per MCS                            kernels2d.cu:95,104,179-182
 *   - MINSTD generator (compat / known-answer only)       random.cu:5-12
 * and make it deterministic (the reference races on its cell tables):
 *   - counter-based Philox4x32-10 (Salmon et al., SC'11) keyed by
 *     (seed; sub-block, MCS, phase, attempt): word 0 the position, 1 the
 *     neighbour, 2 the acceptance; with c = 1 the same-colour sub-blocks
 *     2j and 2j+1 along x share the even one's block (odd: words 3 and 0);
 *   - cell tables read from a phase-start snapshot, updates applied as
 *     commutative integer sums;
 *   - integer dH in a dyadic fixed-point scale and a host-built acceptance
 *     threshold table thr[k] = floor(exp(-k/(scale*T)) * 2^32).
 */
#ifndef CPM_ORACLE_H
#define CPM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NTYPES 20
#define ORC_MAX_SCALE_LOG2 8

/* Parameter block: settings.h:4-31 restricted to the hot-path fields. */
typedef struct {
    int32_t J[ORC_NTYPES][ORC_NTYPES];   /* adhesion (int16 range), settings.h:20 */
    int32_t S[ORC_NTYPES][ORC_NTYPES];   /* stickiness, settings.h:22 */
    double lam_area[ORC_NTYPES];         /* settings.h:11 */
    int32_t target_area[ORC_NTYPES];     /* settings.h:12 */
    double lam_perim[ORC_NTYPES];        /* settings.h:17 */
    int32_t target_perim[ORC_NTYPES];    /* settings.h:18 */
    double temperature;                  /* settings.h:30 */
    int32_t fixed[ORC_NTYPES];           /* settings.h:24: unmovable types */
    double lam_chem[ORC_NTYPES];         /* settings.h: chemotaxis (cpm.cu:155) */
    const int32_t* field_q;              /* field [d][L^d] quantized to 2^-16, or NULL */
    double lam_act[ORC_NTYPES];          /* settings.h: actLambda (cpm.cu:131-135) */
    int32_t act_max[ORC_NTYPES];         /* settings.h: actMax */
    uint32_t* act;                       /* act lattice [L^d] (MCS of the last copy into
                                            each site), updated by sweeps; or NULL */
    double lam_pers[ORC_NTYPES];         /* settings.h: persistenceLambda (cpm.cu:112-115) */
    const int32_t* pref_q;               /* preferred directions [(ncells+1)*3] in 2^-16, or NULL */
} orc_params;

/* Persistence (kernels3d.cu:327-344; 3D only, as in the reference): a copy
 * adds -lambda (d . p), d the unit vector from source to target, p the
 * preferred direction of the source cell (of the target cell when the source
 * is medium), lambda of the source type (of the target type when the source
 * is medium).  In fixed point exactly as chemotaxis, with the field replaced
 * by p quantized to 2^-16: E = (m sum_i o_i p_q,i + 2^23) >> 24.  The scale
 * is 2^8. */

/* Preferred-direction update after each run chunk (kernels.cu:158-212,
 * cpm.cu:487-514): centroids c = anchor + csum / area; the ring of H past
 * centroids takes the current ones at *active, *active advances, and the
 * entry now at *active (H chunks old) is the previous position; for each
 * live cell, u = wrap(c - prev) / |wrap(c - prev)| (no update when zero),
 * p = normalize(u (1 - D) + p D) with D the persistence diffusion of the
 * cell's type.  ctype[cell] is the cell's type.  pref and ring are doubles
 * [(ncells+1)*3] and [H*(ncells+1)*3]; pref_q receives llround(p 2^16). */
void orc_pref_centroids(int L, int d, int ncells, const int32_t* area, const int32_t* anchor,
                        const int64_t* csum, double* out);
void orc_update_pref_dirs(int L, int d, int ncells, const int32_t* area, const int32_t* anchor,
                          const int64_t* csum, const int32_t* ctype, const double* diffusion,
                          double* pref, double* ring, int H, int* active, int32_t* pref_q);
/* Initial preferred directions drawn from the seed (cpm.cu:225-246 draws
 * unit vectors from random_device): uniform on the sphere by rejection, with
 * basic arithmetic only so that every compiler gives the same bits.  For
 * cell c, attempt j = 0, 1, ...: Philox4x32-10 with counter
 * (c, 0x50524546, j, 0) gives x, y, z = w / 2^31 - 1 for w0..w2; the first
 * attempt with 0 < s = x^2 + y^2 + z^2 <= 1 gives p = (x, y, z) / sqrt(s). */
void orc_init_pref_dirs(int ncells, uint64_t seed, double* pref);

/* Act term (kernels2d.cu:312-343,386-397,417-473,532-534) in fixed point.
 * Each site's act value is v = min(max(actMax + act[site] - tick, 0), actMax)
 * (tick = the MCS counter).  G(site, cell) is the geometric mean of v over
 * the site and its neighbours holding the same cell value (0 for medium).
 * A copy adds -lambda/actMax (G_source - G_target), lambda and actMax of the
 * target type (of the source type when the target is medium); the source
 * uses its type's actMax, the target its own.  An accepted copy sets
 * act[target] = tick.  In fixed point:
 *   LG[v] = llround(log2(v) 2^24) (v >= 1), mean = sum LG / count,
 *   G = 2^(mean / 2^24) by a 256-entry table and a second-order term,
 *   held in units of 2^-16 (orc_act_geo_mean);
 *   E = (-mA (Gs - Gt) + 2^31) >> 32, mA = llround(lambda scale / actMax 2^16).
 * With an act term the scale is 2^8.  It needs actMax in [1, 65535],
 * |lambda scale / actMax| < 2^14.  The source's neighbourhood reaches two
 * sites from the target, where with shared roles and c = 1 other targets of
 * the same colour phase lie: such phases are made in two passes, every
 * attempt decided against the phase-start lattice, then the accepted ones
 * applied.  With cell_sync the decide pass is also the claim pass (the
 * accepted proposals claim their cells) and the apply pass commits the
 * decided flips whose sub-block holds both claims. */
#define ORC_ACT_MAX 65535
int64_t orc_act_log2(int v);                          /* LG[v] */
int64_t orc_act_geo_mean(int64_t sum_lg, int count);  /* G in 2^-16 units */
int64_t orc_act_multiplier(double lambda, int act_max, int scale);
/* The act term of copying site s into site t at MCS tick, units of 1/scale. */
int64_t orc_act_energy(int L, int d, const uint32_t* lat, const orc_params* p, int scale,
                       int64_t t, int64_t s, uint32_t tick);

/* Chemotaxis (kernels2d.cu:325-334, kernels3d.cu:314-325) in fixed point.
 * A copy from source s to target t = s - o (o = source - target, one of the
 * 8 / 26 neighbour offsets) adds -lambda (d . f(s)) with d = (t - s) / |t - s|
 * = -o / |o|, i.e. +lambda (o . f(s)) / sqrt(k), k = |o|^2 in {1, 2, 3}.
 * lambda is that of the source type, or of the target type when the source
 * is medium.  In units of 1/scale:
 *   m[type][k] = llround(lambda * scale * 2^8 / sqrt(k)),
 *   g = sum_i o_i f_q,i(s)   (f_q = llround(f * 2^16)),
 *   E = (m g + 2^23) >> 24   (arithmetic shift: rounds half up).
 * With a chemotaxis term the scale is 2^8.  Fields must satisfy |f| < 2^15
 * and |lambda| <= 2^13 so that m g fits in 63 bits. */
#define ORC_FIELD_FRAC 16
#define ORC_CHEM_FRAC 8
int64_t orc_chem_multiplier(double lambda, int scale, int k);
int orc_quantize_field(int64_t n, const double* f, int32_t* out);

/* Generators. */
int32_t orc_minstd_next_f64(int32_t seed);      /* random.cu:5-12, double formula */
uint32_t orc_minstd_next_int(uint32_t x);       /* 16807 x mod (2^31-1), integer */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Fixed-point scale: smallest 2^k (k <= 8) making every lambda integral, or -1. */
int orc_energy_scale(const orc_params* p);
/* Acceptance table; returns its length n (thr[k] for 0 <= k < n; thr[k>=n] = 0). */
int orc_threshold_table(double temperature, int scale, uint32_t* out, int cap);

/* Full recount of the cell tables from a lattice (kernels2d.cu:5-41).
 * area/perim: [ncells+1]; anchor: [(ncells+1)*3] site coordinates of each
 * cell's first site in raster order; csum: [(ncells+1)*3] sums of
 * coordinates unwrapped about the anchor.  L must be a power of two (the
 * periodic wrap is a mask, as in random.cu:25).  Returns 0, -1 if a spin
 * exceeds ncells or a type exceeds 19, -2 for a bad L or d. */
int orc_tally(int L, int d, int ncells, const uint32_t* lat, int32_t* area,
              int32_t* perim, int32_t* anchor, int64_t* csum);

/* Run nmcs Monte Carlo steps starting at MCS counter mcs0.
 * c = positions_per_checkerboard, u = updates_per_checkerboard_switch.
 * nthreads > 1 uses OpenMP over the sub-blocks of one colour phase (results
 * are identical for every thread count).  Tables must hold a valid tally on
 * entry and are kept exact.  Returns 0, or a negative error code. */
int orc_sweep(int L, int d, int ncells, uint32_t* lat, int32_t* area, int32_t* perim,
              const int32_t* anchor, int64_t* csum, const orc_params* p, int c, int u,
              uint64_t seed, uint32_t mcs0, int nmcs, int nthreads);

/* orc_sweep with cell_sync (kernels.cu:123-141 restated deterministically,
 * u = 1 only): in each colour phase every accepted proposal claims its
 * non-medium cells for the lowest sub-block index, and a flip commits only
 * if its sub-block holds both claims (at most one flip per cell per phase).
 * Returns -7 for u != 1. */
int orc_sweep_sync(int L, int d, int ncells, uint32_t* lat, int32_t* area, int32_t* perim,
                   const int32_t* anchor, int64_t* csum, const orc_params* p, int c, int u,
                   uint64_t seed, uint32_t mcs0, int nmcs, int nthreads, int cell_sync);

/* The run() schedule (kernels2d.cu:72-106,179-243, cpm.cu:405):
 * shared = 1: regions of 2c sites, K = 2 colours per axis, the target drawn in
 * the sub-block and the source a neighbour; shared = 0 (the reference's
 * default): regions of ppt = positions_per_thread sites, K = ppt / c colours
 * per axis, the source drawn in the sub-block and the target a neighbour.
 * One MCS is c^d / u rounds of K^d colour phases (cbX outermost); a colour
 * phase covers sub-block (region r) K + colour of every region.  Without
 * shared roles an attempt writes one site outside its sub-block and reads
 * two, so the regions of a colour are processed in S^d groups (r mod S), S
 * the smallest power of two with S ppt - c >= 3 (at most L / ppt): the
 * reference runs them together and races when ppt - c < 3.  The Philox
 * phase index is (round K^d + colour) S^d + group.  orc_schedule returns K
 * and S (or -3 for a bad c / ppt). */
int orc_schedule(int L, int d, int c, int shared, int ppt, int* K, int* S);
int orc_sweep_ex(int L, int d, int ncells, uint32_t* lat, int32_t* area, int32_t* perim,
                 const int32_t* anchor, int64_t* csum, const orc_params* p, int c, int u,
                 uint64_t seed, uint32_t mcs0, int nmcs, int nthreads, int cell_sync, int shared,
                 int ppt);

/* orc_sweep_ex counting, once per attempt, into counts[4]: attempts whose
 * source and target differ, evaluated (after the fixed-type rules),
 * accepted, committed. */
int orc_sweep_counts(int L, int d, int ncells, uint32_t* lat, int32_t* area, int32_t* perim,
                     const int32_t* anchor, int64_t* csum, const orc_params* p, int c, int u,
                     uint64_t seed, uint32_t mcs0, int nmcs, int nthreads, int cell_sync, int shared,
                     int ppt, int64_t* counts);

/* Live-table checkerboard (statistics only): orc_sweep_ex's schedule,
 * stream and decisions, each phase's attempts made one at a time in a
 * per-phase pseudo-random order, every attempt reading the area / perimeter
 * tables as the earlier attempts of its phase left them.  No cell_sync,
 * reference stream or split act path (-10). */
int orc_sweep_live(int L, int d, int ncells, uint32_t* lat, int32_t* area, int32_t* perim,
                   const int32_t* anchor, int64_t* csum, const orc_params* p, int c, int u,
                   uint64_t seed, uint32_t mcs0, int nmcs, int shared, int ppt);

/* orc_sweep_ex with the reference's random stream in place of the counter
 * stream (one MINSTD state per region of R sites, a reference thread; see
 * cpm_oracle.c), one call = one reference launch: the states are seeded from
 * the regions' first-site indices + (int32) seed at entry.  Acceptance:
 * orc_ref_threshold_table.  shared = 1 with ppt > 2c: the reference's global
 * blocks of tpb threads per axis, one state per thread of ppt sites (-9 when
 * 2c does not divide ppt or tpb ppt does not divide L).  With
 * trace != NULL the run is serial and every attempt made appends (region,
 * target site, source site, acceptance draw or -1) to trace[4 * trace_cap]. */
int orc_sweep_refstream(int L, int d, int ncells, uint32_t* lat, int32_t* area, int32_t* perim,
                        const int32_t* anchor, int64_t* csum, const orc_params* p, int c, int u,
                        uint64_t seed, uint32_t mcs0, int nmcs, int nthreads, int cell_sync,
                        int shared, int ppt, int64_t* trace, int64_t trace_cap, int64_t* trace_n,
                        int tpb);
int orc_ref_threshold_table(double temperature, int scale, uint32_t* out, int cap);

/* Serial random-sequential CPM with live tables (statistical comparisons
 * only; not bit-comparable with orc_sweep). */
int orc_serial_sweep(int L, int d, int ncells, uint32_t* lat, int32_t* area, int32_t* perim,
                     const int32_t* anchor, int64_t* csum, const orc_params* p, uint64_t seed,
                     uint32_t mcs0, int nmcs);

/* Scaled integer dH of copying site s into site t (both linear indices) with
 * the given tables (used to pin dH against H(after)-H(before)). */
int64_t orc_delta_h(int L, int d, const uint32_t* lat, const int32_t* area,
                    const int32_t* perim, const orc_params* p, int scale, int64_t t,
                    int64_t s);

/* Full Hamiltonian in FP64 (stickiness excluded; J taken as given, so use a
 * symmetric J for an energy):
 *   H = sum_{unordered neighbour pairs, s_i != s_j} J[t_i][t_j]
 *     + sum_{c>0} lam_A (A_c - A0)^2 + sum_{c>0} lam_P (P_c - P0)^2 */
double orc_hamiltonian(int L, int d, int ncells, const uint32_t* lat, const orc_params* p);

/* Sufficient statistics of H used by dH/dtheta (FP64-exact integers):
 *   pairs[a][b] (a <= b, upper triangle of a 20x20 int64 array): number of
 *       unordered unlike-cell neighbour pairs with types {a, b};
 *   per type t: sa2[t] = sum (A-A0)^2, sa1[t] = sum (A-A0),
 *               sp2[t] = sum (P-P0)^2, sp1[t] = sum (P-P0)  over cells c>0 of type t.
 * ctype[c] gives the type of cell c (cells with area 0 are skipped). */
int orc_energy_stats(int L, int d, int ncells, const uint32_t* lat, const orc_params* p,
                     int64_t* pairs, int64_t* sa2, int64_t* sa1, int64_t* sp2,
                     int64_t* sp1);

/* Centroid readback: anchor + csum/area wrapped into [0, L); NaN for area 0. */
void orc_centroids(int L, int d, int ncells, const int32_t* area, const int32_t* anchor,
                   const int64_t* csum, double* out /* [(ncells+1)*d] */);

#ifdef __cplusplus
}
#endif
#endif
