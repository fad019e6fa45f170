// cpm_device.h -- shared host/device declarations of the MI355X CPM engine.
//
// Lattice encoding follows the reference (src/kernels.cu:44-46, cpm.cu:163):
// a site holds cell | type << 24; cell 0 is the medium.  Sites are stored
// row-major with x fastest (kernels.cu:20-42); L is a power of two.
#pragma once

#include <stdint.h>

#include <hip/hip_runtime.h>

namespace gcpm {

constexpr int kNTypes = 20;  // settings.h:4
constexpr uint32_t kSpinMask = 0xFFFFFFu;

// Hamiltonian parameters in fixed point: every entry is the reference value
// multiplied by the engine's dyadic scale, so dH is an exact integer.  Only
// the nt types present in the lattice are stored (nt = 1 + largest type at
// push), as 2 nt^2 + 4 nt int32 words that kernels stage verbatim:
//   J  [a * nt + b]  adhesion * scale                      (settings.h:20)
//   JS [a * nt + b]  (adhesion + stickiness) * scale        (settings.h:22)
//   la [a]           lambda_area * scale                    (settings.h:11)
//   a0 [a]           target area                            (settings.h:12)
//   lp [a]           lambda_perimeter * scale               (settings.h:17)
//   p0 [a]           target perimeter                       (settings.h:18)
// When at most four types are present the block is also kept padded to
// four types (s4: J[4][4], JS[4][4], la[4], a0[4], lp[4], p0[4]), which the
// sweep reads with uniform (scalar) loads instead of staging it.
constexpr int kParamWords = 2 * kNTypes * kNTypes + 4 * kNTypes;
constexpr int kSmallTypes = 4;
constexpr int kSmallWords = 2 * kSmallTypes * kSmallTypes + 4 * kSmallTypes;
struct DevParams {
    int32_t w[kParamWords];
    int32_t s4[kSmallWords];
    // Chemotaxis multipliers llround(lambda * scale * 2^8 / sqrt(k)) for the
    // neighbour classes k = |o|^2 = 1, 2, 3 (index 0 unused).
    int64_t chem_m[kNTypes][4];
    // Act: llround(lambda scale / actMax 2^16) and actMax per type.
    int64_t act_m[kNTypes];
    int32_t act_max[kNTypes];
    // Persistence multipliers, the chemotaxis formula with lambda_persistence.
    int64_t pers_m[kNTypes][4];
};

__host__ __device__ inline int param_words(int nt) { return 2 * nt * nt + 4 * nt; }

// Indexing of a compact parameter block of nt types.
struct ParamView {
    const int32_t* w;
    int nt, n2;
    __device__ __forceinline__ int32_t J(int a, int b) const { return w[a * nt + b]; }
    __device__ __forceinline__ int32_t JS(int a, int b) const { return w[n2 + a * nt + b]; }
    __device__ __forceinline__ int32_t la(int a) const { return w[2 * n2 + a]; }
    __device__ __forceinline__ int32_t a0(int a) const { return w[2 * n2 + nt + a]; }
    __device__ __forceinline__ int32_t lp(int a) const { return w[2 * n2 + 2 * nt + a]; }
    __device__ __forceinline__ int32_t p0(int a) const { return w[2 * n2 + 3 * nt + a]; }
};

// Constants of one sweep (one run() call).
//
// Cell tables are packed int64 rows: perimeter * 2^32 + area.  Colour phase
// g (counted from the last settle) reads tab[g & 1], its start-of-phase
// snapshot; every accepted flip adds its delta to W = tab[(g + 1) & 1] and
// to del[g & 1]; and the phase also folds the previous phase's deltas
// del[(g + 1) & 1] into W (then clears them).  W is never read in a phase,
// so after it W holds the end-of-phase state with no extra launch or
// barrier.  A settle after the last phase folds the last deltas into the
// lagging base, leaving both bases equal and both delta buffers zero.
//
// The per-phase kernels may run on a 16-bit copy of the lattice (lat16),
// made when every site fits: cell | type << idbits with idbits = 16 - the
// bits of the largest type, cells up to 2^idbits - 1.  The map is one to
// one on the sites present, so equal sites stay equal and every result is
// that of the 32-bit lattice (idbits = 24); it halves the bytes each phase
// moves.
struct SweepConsts {
    uint32_t* lat;
    uint16_t* lat16;        // the 16-bit lattice, or null
    uint32_t idbits;        // a site holds cell | type << idbits
    uint32_t idmask;        // (1 << idbits) - 1
    int64_t* tab[2];
    int64_t* del[2];
    const DevParams* prm;
    const uint32_t* thr;    // acceptance thresholds, thr[k] for 0 <= k < nthr
    int nthr;
    int ncells;
    int lgL;                // log2 L
    int lc;                 // log2 positions_per_checkerboard
    int u;                  // updates_per_checkerboard_switch
    int rounds;             // c^d / u colour rounds per MCS
    uint32_t k0, k1;        // Philox key (seed)
    int nt;                 // types present (pair-table stride)
    uint32_t fixed_mask;    // bit t set: type t is fixed (settings.h:24)
    int roles;              // 0 shared: target drawn in the sub-block; 1: source drawn
    int K, S;               // colours and groups per axis
    int lstride;            // log2 (S K)
    int phases_per_round;   // K^d S^d
    int chem;               // chemotaxis term on
    const int32_t* field;   // field [d][L^d], quantized to 2^-16 (or null)
    int pers;               // persistence term on (3D)
    const int32_t* pref;    // preferred directions [(n+1)*3], llround(p 2^16)
    int act;                // act term on
    uint32_t* act_lat;      // act lattice [L^d]: MCS of the last copy into each site
    const int64_t* act_tab; // 2^(j/256) 2^30 [256], then llround(log2(v) 2^24) [v]
    int64_t act_ln2q;       // llround(ln 2 2^30)
    // Reference stream (cpm_set_rng): one MINSTD state per region of K
    // sub-blocks per axis (a reference thread), rstate[(rz << 2 lr) | (ry <<
    // lr) | rx]; thr then holds the 31-bit acceptance table.
    int refstream;
    int32_t* rstate;
    int lk, lr;             // log2 K, log2 (reference threads per axis)
    // The reference's shared mode with positions_per_thread > 2c: each
    // reference thread owns 2^lg regions of 2c sites per axis, threads in
    // blocks of 2^ld (threads_per_block) per axis (0, 0 otherwise).
    int lg, ld;
    // Attempt counters (cpm_set_counters), or null: attempts whose source and
    // target differ, attempts evaluated (after the fixed-type rules),
    // proposals accepted, flips committed.
    unsigned long long* counters;
};

// One colour phase per launch.
struct PhaseArgs {
    SweepConsts k;
    // Active sub-blocks: ((q mod 2^lh) << lstride) + cx per axis (sub-block
    // units; q the thread's index per axis): group * K + colour.
    int cx, cy, cz;
    int lstride;            // log2 (S K): sub-blocks between active ones
    uint32_t phase;         // Philox phase index (round K^d + colour) S^d + group
    uint32_t mcs;           // Monte Carlo step counter
    int g;                  // phase counter since the last settle (its parity selects buffers)
    int lh;                 // log2 (same-colour sub-blocks per axis)
    int nsub;               // active sub-blocks in this phase
    int box3;               // 3D, c = 1: stage each block's box in LDS
    int sync_pass;          // 0 off; cell_sync: 1 claim, 2 commit; split phase: 3 decide, 4 apply;
                            // both: 5 decide + claim, 6 apply + claim check
    uint64_t* claims;       // cell_sync: per-cell claim keys
    uint32_t sync_tag;      // cell_sync: phase tag (grows across runs)
    uint8_t* decide;        // split phase: one decision per sub-block of the phase
    int form;               // kForm* bits: which per-phase kernel runs
    int att0, natt;         // attempt indices att0 .. att0 + natt - 1 of each sub-block
    int gsx, gsy, gsz;      // reference stream with lg > 0: the phase's block slice per axis
};

// Per-phase kernel forms (results never depend on them).
constexpr int kFormSmall = 1;    // read at most four types' parameters without LDS staging
constexpr int kFormCompact = 2;  // u = 1: queue the attempts whose source and target differ
constexpr int kFormFull = 4;     // compact: the draw loads (and queues) the whole neighbourhood
constexpr int kFormAtt2 = 8;     // compact: each thread draws 2 attempts
constexpr int kFormAtt4 = 16;    // compact: each thread draws 4 attempts
constexpr int kFormAtt8 = 32;    // compact: each thread draws 8 attempts
__host__ __device__ inline int form_att(int form) {
    return (form & kFormAtt8) ? 8 : (form & kFormAtt4) ? 4 : (form & kFormAtt2) ? 2 : 1;
}
__host__ __device__ inline int att_form(int att) {
    return att == 8 ? kFormAtt8 : att == 4 ? kFormAtt4 : att == 2 ? kFormAtt2 : 0;
}

// A whole run in one cooperative launch, phases separated by grid barriers.
struct PersistArgs {
    SweepConsts k;
    uint32_t mcs0;          // first MCS counter
    int nmcs;               // MCS in this launch
    int lh;                 // log2 (same-colour sub-blocks per axis)
    unsigned int* bar;      // grid barrier {arrivals, generation, abandoned-wait flag}
    int ltx, lty;           // log2 LDS tile edge (2D tiled kernel)
    int lntx;               // log2 tiles per row
};

// A whole inner_iterations chunk in one workgroup with the lattice in LDS.
struct SmallArgs {
    SweepConsts k;
    uint32_t mcs0;          // first MCS counter
    int nmcs;               // MCS in this launch
    int lh;                 // log2 (same-colour sub-blocks per axis)
    int thr_lds;            // 1: the acceptance thresholds are staged in LDS too
};

// The resident form (2D, u = 1): one cooperative block per tile of 2^lte x
// 2^lte sites, 2^lnt tiles per axis, for one launch.  A
// tile keeps its sites, a one-site halo and the rows of the cells it holds
// in LDS; neighbouring tiles exchange border sites and the deltas of the
// cells both hold through tagged mailbox words.
constexpr int kResSlots = 254;   // cells one tile holds at most (slots 254, 255: medium, none)
constexpr int kResDirCap = 64;   // cells one tile shares with one neighbour at most
__host__ __device__ inline int res_mail_stride(int te) { return te + kResDirCap; }
// Why a chunk failed the guard (bits of ResidentArgs::bad): a cell's widened
// box spans more than two tiles per axis; a tile holds more than kResSlots
// cells; it shares more than kResDirCap with one neighbour; an attempt met a
// cell its tile does not hold.
constexpr unsigned int kResGuardWide = 1, kResGuardSlots = 2, kResGuardShared = 4, kResGuardMiss = 8;
// Sites a cell's box is widened by (beyond the halo) for its growth within a
// chunk: a quarter tile.
__host__ __device__ inline int res_margin(int te) { return te / 4; }
struct ResidentArgs {
    SweepConsts k;
    uint32_t mcs0;
    int nmcs;
    int lte, lnt;
    int gen;                   // 0: c = 1 with shared roles (pairs of sites); 1: any schedule with u = 1
    unsigned long long* mail;  // [tiles][2][8][res_mail_stride(2^lte)] tagged words, 0 on entry
    const uint32_t* holders;   // [n + 1] the tiles holding each cell (0: none)
    unsigned int* err;         // abandoned-wait flag, 0 on entry
    unsigned int* bad;         // guard flag: a cell outside its tiles, too many cells
};

// One replica of a batched run.
struct ReplicaDesc {
    SweepConsts k;
    uint64_t* claims;       // cell_sync claim keys of this replica
    uint8_t* decide;        // split phases: one decision byte per sub-block
    uint32_t mcs0;          // this replica's MCS counter at the start
    int lh;                 // log2 (same-colour sub-blocks per axis)
    int block0, nblocks;    // its block group in the cooperative grid
};

// Batched replicas in one cooperative launch (device-resident descriptors).
struct BatchArgs {
    const ReplicaDesc* reps;
    const int* rep_of;      // block -> replica
    unsigned int* bars;     // 2 words per replica {arrivals, generation}
    unsigned int* err;      // abandoned-wait flag of the launch
    int nmcs;               // MCS per replica in this launch
};

// Batched replicas, one launch per colour phase: block b serves replica
// rep_of[b]; every replica has the same c and u, so phases stay in step.
struct BatchPhaseArgs {
    const ReplicaDesc* reps;
    const int* rep_of;
    int cx, cy, cz;         // as PhaseArgs
    int lstride;
    uint32_t phase;         // round * 2^d + colour
    uint32_t m;             // MCS offset within this run (replica MCS = mcs0 + m + *mbase)
    const uint32_t* mbase;  // graph replay: device MCS offset of the replayed block of MCS (or null)
    int g;                  // phase counter since the last settle (parity)
    int sync_pass;          // as PhaseArgs (0 off; 1/2 cell_sync; 3/4 split; 5/6 both)
    uint32_t sync_tag;      // cell_sync: tag (above every replica's earlier tags)
    int att0, natt;         // attempt indices of this launch
    int form;               // kForm* bits (compact, small, attempts per thread)
    int narrow;             // every replica on its 16-bit lattice (else all on 32-bit)
    int refstream;          // every replica draws from its reference stream (not compacted)
    int gsx, gsy, gsz;      // as PhaseArgs
};

// Host-side launchers (kernels.hip).
// kernel_id (optional) receives the kernel used; phase_kernel_name names it.
hipError_t launch_phase(int d, const PhaseArgs& a, hipStream_t s, hipEvent_t e0, hipEvent_t e1,
                        int* kernel_id = nullptr);
const char* phase_kernel_name(int id);
// Cooperative persistent launches.  lds2d selects the 2D LDS-tiled kernel.
// max_blocks returns the co-resident block limit for the chosen kernel.
hipError_t persist_max_blocks(int d, bool lds2d, size_t lds_bytes, int* max_blocks);
hipError_t launch_persist(int d, bool lds2d, const PersistArgs& a, int blocks, size_t lds_bytes,
                          hipStream_t s);
// The single-workgroup form: dynamic LDS bytes it needs, and its launch
// (one block; e0/e1 or null are recorded at the dispatch's start and end).
size_t small_lds_bytes(int d, int lgL, int ncells, bool narrow);
hipError_t launch_small(int d, bool narrow, const SmallArgs& a, hipStream_t s, hipEvent_t e0, hipEvent_t e1);
// The resident form: dynamic LDS bytes (the cell map), co-resident block
// limit and launch; var selects the kernel: 0 the lean attempt path (at
// most four types, no fixed types, chemotaxis or counters, a threshold
// table of at most kResThrLean entries) for c = 1 with shared roles, 3 the
// lean attempt path for any other schedule, 1 the general one, 2 the
// general one with the chemotaxis term.
constexpr int kResThrLean = 1024;
size_t resident_lds_bytes(int ncells);
hipError_t resident_max_blocks(bool narrow, int lte, int var, int ncells, int* max_blocks);
hipError_t launch_resident(bool narrow, int var, const ResidentArgs& a, int blocks, hipStream_t s);
// Before a chunk: the boxes of the current lattice's cells into bbox, and
// the holder word of each cell; *bad |= 1 if a cell's widened box spans
// more than two tiles per axis.
// The copies and fills before a resident launch (one kernel).
struct ResPrep {
    const void* lat;  // the current lattice (16- or 32-bit), lat_bytes a multiple of 16
    void* lat_backup;
    size_t lat_bytes;
    unsigned long long* mail;  // nmail words zeroed (nmail even)
    size_t nmail;
    const int64_t* tab;  // rows copied to tab_backup; bbox rows set empty
    int64_t* tab_backup;
    int32_t* bbox;
    int rows;
    unsigned int* flags;             // the two error words, zeroed
    unsigned long long* counters;    // if set: counters[4..7] = counters[0..3]
};
hipError_t launch_res_prep(const ResPrep& q, hipStream_t s);
// Boxes (set empty beforehand), holder words and the box guard of a resident launch.
hipError_t launch_resident_guard(bool narrow, const SweepConsts& k, const int32_t* anchor, int32_t* bbox,
                                 int lte, int lnt, uint32_t* holders, unsigned int* bad, hipStream_t s);
// *p = v / *p += v (one thread: nodes of a captured graph, no memset nodes)
hipError_t launch_set_u32(uint32_t* p, uint32_t v, hipStream_t s);
hipError_t launch_add_u32(uint32_t* p, uint32_t v, hipStream_t s);
hipError_t batch_max_blocks(int d, int* max_blocks);
hipError_t launch_batch(int d, const BatchArgs& a, int blocks, hipStream_t s);
// e0/e1 (or null) are recorded at the dispatch's start and end; name (or
// null) receives the kernel's name as profilers show it.
hipError_t launch_batch_phase(int d, const BatchPhaseArgs& a, int blocks, hipStream_t s,
                              hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr, char* name = nullptr,
                              size_t name_len = 0);
hipError_t launch_tally(int d, const uint32_t* lat, int lgL, int ncells, int64_t* tab,
                        uint32_t* first, int32_t* anchor, hipStream_t s);
// Settle after G phases of a run (only the parity of G matters).
hipError_t launch_settle(int64_t* const tab[2], int64_t* const del[2], int G, int ncells,
                         hipStream_t s);
// Centroid sums of the current lattice about the push-time anchors
// (csum must be zero on entry).
// Copy the 32-bit lattice into its 16-bit form (to_narrow) or back; sites % 4 == 0.
// Seed every region's reference stream for one reference launch:
// rstate[r] = MINSTD(index of the region's first site + seed32).
hipError_t launch_ref_seed(int d, const SweepConsts& k, uint32_t seed32, hipStream_t s);
hipError_t launch_convert(uint32_t* wide, uint16_t* narrow, size_t sites, int idbits, bool to_narrow,
                          hipStream_t s);
hipError_t launch_csum(int d, const uint32_t* lat, int lgL, const int32_t* anchor, int64_t* csum,
                       hipStream_t s);
// Energy statistics of one engine in a batched statistics launch: pair
// counts over its lattice (16- or 32-bit sites) and per-type moments over
// its settled cell table.  out[20*20 + 20*5] must be zero on entry.  Blocks
// [pair0, cell0) of the launch count its pairs (stats_pair_blocks of them),
// [cell0, cell0 + stats_cell_blocks) its cell moments.
struct StatsDesc {
    const void* lat;
    int narrow;             // lat holds 16-bit sites (cell | type << idbits); equal over a launch
    int lgL;
    uint32_t idbits;
    int nt;                 // types present (at most 4: packed counters)
    const int64_t* tab;     // settled table rows (perimeter << 32 | area)
    const int32_t* ctype;   // [n+1] type of each cell (set at push)
    int ncells;
    unsigned long long* out;
    int pair0, cell0;
};
constexpr int kStatsWords = 20 * 20 + 20 * 5;   // words of one engine's statistics
constexpr int kStatsArgEngines = 32;             // engines per argument-borne launch
constexpr int kStatsScratchEngines = 64;         // engines of the allocation-free entry point
constexpr int kMaxDevices = 64;
// Descriptors of up to kStatsArgEngines engines by value (a kernel argument):
// engine r owns the launch's blocks [d[r].pair0, end[r]).
struct StatsArgs {
    StatsDesc d[kStatsArgEngines];
    int end[kStatsArgEngines];
    int n;
};
// The per-device address of the allocation-free entry point's scratch
// (a file-scope device array of kStatsScratchEngines * kStatsWords words).
hipError_t stats_scratch(unsigned long long** out);
// n words at p := 0 (a fill kernel, not a memset node).
hipError_t launch_fill_zero(unsigned long long* p, size_t n, hipStream_t s);
// t3: every engine of the launch is 2D, 16-bit, L >= 8, at most three types
hipError_t launch_stats_args(int d, bool narrow, bool t3, const StatsArgs& a, int blocks, hipStream_t s);
int stats_pair_blocks(int d, int lgL);
int stats_cell_blocks(int ncells);
// every engine of one launch has the same dimension and site width
hipError_t launch_stats_batch(int d, bool narrow, bool t3, const StatsDesc* descs, const int* eng_of, int blocks,
                              hipStream_t s, hipEvent_t e0, hipEvent_t e1);

constexpr int kBlock = 256;  // threads per workgroup of the sweep kernels

}  // namespace gcpm
