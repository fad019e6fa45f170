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
// multiplied by the engine's dyadic scale, so dH is an exact integer.
struct DevParams {
    int32_t J[kNTypes][kNTypes];   // adhesion * scale            (settings.h:20)
    int32_t JS[kNTypes][kNTypes];  // (adhesion + stickiness) * scale (settings.h:22)
    int32_t la[kNTypes];           // lambda_area * scale         (settings.h:11)
    int32_t a0[kNTypes];           // target area                 (settings.h:12)
    int32_t lp[kNTypes];           // lambda_perimeter * scale    (settings.h:17)
    int32_t p0[kNTypes];           // target perimeter            (settings.h:18)
};

// Arguments of one colour phase of the checkerboard sweep.
//
// Cell tables are packed int64 rows: perimeter * 2^32 + area.  The phase
// reads R (its start-of-phase snapshot), adds every accepted flip's delta to
// W and to Dcur, and first folds the previous phase's deltas Dprev into W
// (then clears Dprev).  With R/W and Dcur/Dprev swapping each phase, W
// equals the end-of-phase state without any extra launch or barrier.
struct PhaseArgs {
    uint32_t* lat;
    const int64_t* R;
    int64_t* W;
    int64_t* Dcur;
    int64_t* Dprev;
    const int32_t* anchor;  // [(n+1)*3]
    int64_t* csum;          // [(n+1)*3]
    const DevParams* prm;
    const uint32_t* thr;    // acceptance thresholds, thr[k] for 0 <= k < nthr
    int nthr;
    int ncells;
    int lgL;                // log2 L
    int lc;                 // log2 positions_per_checkerboard
    int lh;                 // log2 (same-colour sub-blocks per axis) = lgL - lc - 1
    int u;                  // updates_per_checkerboard_switch
    int cx, cy, cz;         // colour parity
    uint32_t phase;         // round * 2^d + colour
    uint32_t mcs;           // Monte Carlo step counter
    uint32_t k0, k1;        // Philox key (seed)
    int nsub;               // active sub-blocks in this phase
};

// Host-side launchers (kernels.hip).
hipError_t launch_phase(int d, const PhaseArgs& a, hipStream_t s, hipEvent_t e0, hipEvent_t e1);
hipError_t launch_tally(int d, const uint32_t* lat, int lgL, int ncells, int64_t* tab,
                        uint32_t* first, int32_t* anchor, int64_t* csum, hipStream_t s);
hipError_t launch_settle(int64_t* tab, int64_t* del, int ncells, hipStream_t s);

}  // namespace gcpm
