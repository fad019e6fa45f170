// engine.cpp -- host engine and C ABI of the MI355X CPM engine.
//
// Mirrors the reference host engine class Cpm (shabaz/gpu-cpm src/cpm.h,
// src/cpm.cu) with per-instance parameters (the reference keeps one global
// __constant__ Settings, kernels.cu:48-53), validated inputs, error codes
// instead of printed CUDA errors (cpm.cu:52-56), and a deterministic,
// resumable schedule keyed by (seed, MCS counter).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <random>
#include <string>
#include <vector>

#include "../../include/gpucpm.h"
#include "cpm_device.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

int hip_fail(hipError_t e, const char* what) {
    return fail(CPM_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(expr)                                  \
    do {                                               \
        hipError_t e_ = (expr);                        \
        if (e_ != hipSuccess) return hip_fail(e_, #expr); \
    } while (0)

bool is_pow2(long v) { return v > 0 && (v & (v - 1)) == 0; }

int ilog2(long v) {
    int k = 0;
    while ((1L << k) < v) k++;
    return k;
}

constexpr int kMaxScaleLog2 = 8;

// How one run is launched.  All three forms give identical results.
struct Plan {
    enum Mode { kPhase, kGlobal, kLds2d, kSmall, kResident } mode = kPhase;
    int blocks = 0;        // cooperative grid size
    int ltile = 0;         // log2 tile edge (kLds2d)
    size_t lds_bytes = 0;  // dynamic LDS per block (kLds2d)
};
constexpr int kMaxThr = 1 << 22;
// The single-workgroup form: dynamic LDS it may use (of the 160 KiB of a
// gfx950 workgroup, leaving room for its static LDS), and the largest number
// of attempts per phase for which it is chosen automatically.
constexpr size_t kSmallLdsBudget = 158 * 1024;
constexpr long kSmallAutoAttempts = 4096;
// Batch graph replay: the largest number of MCS captured in one graph.
constexpr long kGraphMcs = 64;
// Resident form: the shortest launch (MCS) the adaptive launch length goes
// down to after guard failures.
constexpr int64_t kResMinMcs = 4;

// Per-phase kernel options (cpm_set_kernel_options); -1 keeps the default.
// Every option computes the same results.
struct KernelOpts {
    int small = -1;    // at most four types: parameters from scalar loads (default in 3D)
    int compact = -1;  // u = 1: queue the attempts whose sites differ (default on)
    int full = -1;     // compacted queue carries the whole neighbourhood (default off)
    int att = -1;      // attempts drawn per thread by the compacted kernel (1, 2, 4, 8)
    int narrow = -1;   // 16-bit sites when they fit (default on)
    int box3 = -1;     // 3D, c = 1: stage 32-bit boxes in LDS (default off)
};

int pick(int v, int dflt) { return v >= 0 ? v : dflt; }

// Online autotuner of the per-phase kernel form (cpm_set_autotune).  Every
// form computes the same results, so the tuner times a run's own launches:
// while a workload key is unsettled each launch takes the candidate with the
// fewest samples and is bracketed by events; the timings are read when the
// run ends, and once every candidate has kSamples the fastest mean is kept
// for that key.  Samples carry over between runs (interactive runs of one MCS
// settle after a few calls) and between engines: the table is process-wide,
// the pending timings belong to one engine or batch.
struct FormTuner {
    static constexpr int kSamples = 8;
    struct Entry {
        std::vector<int> forms;
        std::vector<double> ms;
        std::vector<int> n;
        int best = -1;  // the chosen form, or -1 while sampling
    };
    struct Pending {
        Entry* e;
        int idx;
        hipEvent_t e0, e1;
        double scale = 1.0;  // the sample is the elapsed time divided by this (e.g. MCS of a run)
    };
    static std::map<uint64_t, Entry>& table() {
        static std::map<uint64_t, Entry> t;
        return t;
    }
    static std::mutex& lock() {
        static std::mutex m;
        return m;
    }
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;
    size_t used = 0;

    Entry& entry(uint64_t key, const std::vector<int>& forms) {
        std::lock_guard<std::mutex> g(lock());
        Entry& e = table()[key];
        if (e.forms.empty()) {
            e.forms = forms;
            e.ms.assign(forms.size(), 0.0);
            e.n.assign(forms.size(), 0);
            if (forms.size() == 1) e.best = forms[0];
        }
        return e;
    }
    // Candidate for the next launch, or -1 when every candidate already has
    // its samples (recorded or pending).
    int pick(const Entry& e) const {
        std::lock_guard<std::mutex> g(lock());
        int idx = -1, least = kSamples;
        for (size_t i = 0; i < e.forms.size(); i++) {
            int c = e.n[i];
            for (const Pending& p : pending)
                if (p.e == &e && p.idx == (int)i) c++;
            if (c < least) {
                least = c;
                idx = (int)i;
            }
        }
        return idx;
    }
    hipError_t events(hipEvent_t* e0, hipEvent_t* e1) {
        while (pool.size() < used + 2) {
            hipEvent_t ev;
            hipError_t err = hipEventCreate(&ev);
            if (err != hipSuccess) return err;
            pool.push_back(ev);
        }
        *e0 = pool[used];
        *e1 = pool[used + 1];
        used += 2;
        return hipSuccess;
    }
    // Read the pending timings (waits for their launches) and settle the
    // entries that have all their samples.
    hipError_t settle() {
        if (pending.empty()) return hipSuccess;
        hipError_t err = hipEventSynchronize(pending.back().e1);
        std::lock_guard<std::mutex> g(lock());
        for (size_t i = 0; i < pending.size() && err == hipSuccess; i++) {
            const Pending& p = pending[i];
            float ms = 0.0f;
            if ((err = hipEventElapsedTime(&ms, p.e0, p.e1)) != hipSuccess) break;
            p.e->ms[p.idx] += ms / p.scale;
            p.e->n[p.idx]++;
        }
        if (err != hipSuccess) {  // drop this run's samples rather than retry them
            pending.clear();
            used = 0;
            return err;
        }
        for (const Pending& p : pending) {
            Entry& e = *p.e;
            if (e.best >= 0) continue;
            bool done = true;
            for (int c : e.n) done = done && c >= kSamples;
            if (!done) continue;
            size_t bi = 0;
            for (size_t i = 1; i < e.forms.size(); i++)
                if (e.ms[i] / e.n[i] < e.ms[bi] / e.n[bi]) bi = i;
            e.best = e.forms[bi];
        }
        pending.clear();
        used = 0;
        return hipSuccess;
    }
    ~FormTuner() {
        for (hipEvent_t e : pool) (void)hipEventDestroy(e);
    }
};

// Philox4x32-10 on the host (Salmon et al., SC'11), for draws made outside
// the sweep (initial preferred directions).
void host_philox(uint32_t x0, uint32_t x1, uint32_t x2, uint32_t x3, uint32_t k0, uint32_t k1,
                 uint32_t w[4]) {
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

// kernels.cu:150-156
double wrap_half(double x, int L) {
    if (x > L / 2) x -= L;
    if (x < -L / 2) x += L;
    return x;
}

}  // namespace

// Host mirror of the hot-path fields of the reference Settings
// (settings.h:4-31); J and stickiness are int16 there.
struct HostSettings {
    double J[CPM_NTYPES][CPM_NTYPES];  // int16 values (or dyadic reals via cpm_set_adhesion_real)
    int16_t S[CPM_NTYPES][CPM_NTYPES];
    double lam_area[CPM_NTYPES];
    int32_t target_area[CPM_NTYPES];
    double lam_perim[CPM_NTYPES];
    int32_t target_perim[CPM_NTYPES];
    double temperature;
    int32_t fixed[CPM_NTYPES];  // settings.h:24
    double chem[CPM_NTYPES];    // chemotaxis lambda (cpm.cu:155)
    double lam_act[CPM_NTYPES];  // actLambda (cpm.cu:131-135)
    int32_t act_max[CPM_NTYPES]; // actMax
    double lam_pers[CPM_NTYPES];   // persistenceLambda (cpm.cu:112-115)
    double pers_diff[CPM_NTYPES];  // persistenceDiffusion
    int32_t pers_time[CPM_NTYPES]; // persistenceTime (stored; the device path does not read it)
};

struct cpm_engine {
    int L = 0, d = 0, n = 0, lgL = 0;
    size_t sites = 0;
    int active = 0;
    bool pushed = false;
    bool params_dirty = true;
    uint64_t seed = 0;
    uint64_t mcs = 0;
    HostSettings st{};

    uint32_t* h_lat = nullptr;  // pinned host mirror
    bool has_field = false;
    std::vector<double> h_field;         // host field [d][sites] (cpm.cu:189-190)
    std::vector<double> h_field_pushed;  // the field as last pushed (pull restores it)
    int32_t* d_field = nullptr;          // device field, quantized to 2^-16
    bool has_act = false;
    uint32_t* h_act = nullptr;           // host act lattice (cpm.cu:193), pinned
    uint32_t* d_act = nullptr;
    int64_t* d_act_tab = nullptr;        // act 2^x table [256] + log2 table [amax + 1]
    int act_tab_cap = 0;
    // Persistence (3D): preferred directions, the centroid history ring of
    // `history` entries (cpm.cu:220-223), cell types.
    int history = 1;
    std::vector<double> pref;            // [(n+1)*3]
    std::vector<double> ring;            // [history][(n+1)*3]
    int ring_active = 0;
    bool ring_valid = false;             // filled at the first persistence run after a push
    bool pref_dirty = true;              // d_pref needs an upload
    int32_t* d_pref = nullptr;           // [(n+1)*3] llround(p 2^16)
    std::vector<int32_t> ctype;          // [n+1] type of each cell at push
    hipStream_t stream = nullptr;
    hipEvent_t ev_ready = nullptr;  // marks this engine's queued work for other streams to wait on
    uint32_t* d_lat = nullptr;
    // 16-bit form of the lattice (cell | type << lat16_bits) for the per-phase
    // kernels; wide_valid / narrow_valid say which copy holds the current
    // state (at least one does).
    uint16_t* d_lat16 = nullptr;
    int lat16_bits = 0;
    // Random stream: 0 the counter stream (Philox), 1 the reference's
    // per-thread MINSTD stream (cpm_set_rng), whose states live in d_rstate.
    int rng = 0;
    int32_t* d_rstate = nullptr;
    size_t rstate_cap = 0;
    uint8_t* d_decide = nullptr;  // split phases (act, shared roles, c = 1): decisions
    size_t decide_cap = 0;
    bool wide_valid = true;
    bool narrow_valid = false;
    int64_t* d_tab[2] = {nullptr, nullptr};
    int64_t* d_del[2] = {nullptr, nullptr};
    int32_t* d_anchor = nullptr;
    int64_t* d_csum = nullptr;
    uint32_t* d_first = nullptr;
    gcpm::DevParams* d_prm = nullptr;
    unsigned int* d_bar = nullptr;  // grid barrier state {arrivals, generation, abandoned-wait flag}
    uint64_t* d_claims = nullptr;           // cell_sync claim keys [n+1]
    uint32_t sync_tag = 0;                  // cell_sync phase tag
    int32_t* d_ctype = nullptr;             // [n+1] cell types at push (energy statistics)
    uint32_t* d_thr = nullptr;
    int thr_cap = 0;
    int nthr = 0;
    int scale = 1;
    int nt = 1;  // types present in the lattice at the last push
    int last_kernel = -1;  // per-phase kernel of the last run (gcpm::phase_kernel_name)
    bool small_narrow = false;  // the last single-workgroup or resident run used 16-bit sites
    int res_var = 0;            // the last resident run's kernel variant (gcpm::launch_resident)
    // Resident form (2D): abandoned-wait flag, guard flag and the cells'
    // holder words; the tiles' mailboxes; per-cell boxes; and the copy of
    // lattice and table a failed chunk is restored from.
    unsigned int* d_res_flags = nullptr;
    unsigned long long* d_res_mail = nullptr;
    size_t res_mail_cap = 0;
    int32_t* d_bbox = nullptr;
    void* d_backup_lat = nullptr;
    int64_t* d_backup_tab = nullptr;
    // attempt counters (cpm_set_counters): differing, evaluated, accepted,
    // committed, then a copy of them that a redone resident chunk restores;
    // attempts drawn are L^d per MCS counted (ctr_mcs_acc, plus the MCS since
    // ctr_mcs0 while on)
    unsigned long long* d_counters = nullptr;
    bool counters_on = false;
    uint64_t ctr_mcs0 = 0, ctr_mcs_acc = 0;
    int64_t res_mcs = 0;            // resident form: MCS per launch after a guard failure (0: the chunk)
    int res_streak = 0;             // launches in a row that held since K last changed
    bool res_auto_off = false;      // the automatic choice no longer takes the form (until a push)
    bool res_small_tiles = false;   // tiles of 32^2 where 64^2 held too many cells (until a push)
    int64_t res_fallbacks = 0;      // launches redone with one launch per phase: the guard
    int64_t res_wait_fallbacks = 0; // ... or an abandoned neighbour wait
    bool autotune = false;  // cpm_set_autotune
    FormTuner tuner;

    // Kernel timing (cpm_timing_begin/end): one event pair per sweep launch.
    bool timing = false;
    std::vector<hipEvent_t> ev;
    size_t ev_used = 0;
    int64_t t_launches = 0;
    int64_t t_updates = 0;
    Plan last_plan;  // launch form of the last run
    int want_form = -1;  // cpm_set_launch: -1 auto, else Plan::Mode
    int want_tile = 0;   // cpm_set_launch: LDS tile edge (0 = heuristic)
    KernelOpts opts;     // cpm_set_kernel_options

    int next_events(hipEvent_t* e0, hipEvent_t* e1) {
        *e0 = *e1 = nullptr;
        if (!timing) return CPM_OK;
        while (ev.size() < ev_used + 2) {
            hipEvent_t e;
            HIP_TRY(hipEventCreate(&e));
            ev.push_back(e);
        }
        *e0 = ev[ev_used];
        *e1 = ev[ev_used + 1];
        ev_used += 2;
        return CPM_OK;
    }

    ~cpm_engine() {
        // Teardown is best effort: errors here have no caller to report to.
        if (stream) (void)hipStreamSynchronize(stream);
        (void)hipFree(d_lat);
        (void)hipFree(d_lat16);
        (void)hipFree(d_rstate);
        (void)hipFree(d_decide);
        for (int i = 0; i < 2; i++) (void)hipFree(d_tab[i]);
        for (int i = 0; i < 2; i++) (void)hipFree(d_del[i]);
        (void)hipFree(d_anchor);
        (void)hipFree(d_csum);
        (void)hipFree(d_first);
        (void)hipFree(d_prm);
        (void)hipFree(d_bar);
        (void)hipFree(d_claims);
        (void)hipFree(d_ctype);
        (void)hipFree(d_thr);
        (void)hipFree(d_field);
        (void)hipFree(d_act);
        (void)hipFree(d_pref);
        (void)hipFree(d_act_tab);
        if (h_act) (void)hipHostFree(h_act);
        if (h_lat) (void)hipHostFree(h_lat);
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        if (ev_ready) (void)hipEventDestroy(ev_ready);
        (void)hipFree(d_res_flags);
        (void)hipFree(d_res_mail);
        (void)hipFree(d_bbox);
        (void)hipFree(d_counters);
        (void)hipFree(d_backup_lat);
        (void)hipFree(d_backup_tab);
        if (stream) (void)hipStreamDestroy(stream);
    }

    bool pers_on() const {
        for (int t = 0; t < CPM_NTYPES; t++)
            if (st.lam_pers[t] != 0.0) return true;
        return false;
    }

    // Initial preferred directions from the seed (cpm.cu:225-246 draws unit
    // vectors from random_device): uniform on the sphere by rejection, basic
    // arithmetic only.  Cell c, attempt j: Philox4x32-10 counter
    // (c, 0x50524546, j, 0) gives x, y, z = w / 2^31 - 1; the first attempt with
    // 0 < s = |(x, y, z)|^2 <= 1 gives p = (x, y, z) / sqrt(s).
    void init_pref_dirs() {
        for (int c = 0; c <= n; c++) {
            for (uint32_t j = 0;; j++) {
                uint32_t w[4];
                host_philox(c, 0x50524546u, j, 0u, (uint32_t)seed, (uint32_t)(seed >> 32), w);
                const double x = (double)w[0] / 2147483648.0 - 1.0;
                const double y = (double)w[1] / 2147483648.0 - 1.0;
                const double z = (double)w[2] / 2147483648.0 - 1.0;
                const double s2 = x * x + y * y + z * z;
                if (s2 > 0.0 && s2 <= 1.0) {
                    const double nn = std::sqrt(s2);
                    pref[3 * c + 0] = x / nn;
                    pref[3 * c + 1] = y / nn;
                    pref[3 * c + 2] = z / nn;
                    break;
                }
            }
        }
        pref_dirty = true;
    }

    // Cell bits of the 16-bit lattice for the current cells and types, or 0
    // when the sites do not fit in 16 bits (or the kernel options turn the
    // 16-bit form off).
    int narrow_bits() const {
        if (opts.narrow == 0) return 0;
        int tb = 0;
        while ((1 << tb) < nt) tb++;
        const int bits = 16 - tb;
        return n <= (1 << bits) - 1 && sites % 4 == 0 ? bits : 0;
    }

    // Make the 32-bit lattice current (after runs on the 16-bit one).
    int ensure_wide() {
        if (!wide_valid) {
            HIP_TRY(gcpm::launch_convert(d_lat, d_lat16, sites, lat16_bits, false, stream));
            wide_valid = true;
        }
        return CPM_OK;
    }

    // Make the 16-bit lattice with `bits` cell bits current.
    int ensure_narrow(int bits) {
        if (narrow_valid && lat16_bits == bits) return CPM_OK;
        if (int rc = ensure_wide()) return rc;
        if (!d_lat16) HIP_TRY(hipMalloc(&d_lat16, sites * 2));
        HIP_TRY(gcpm::launch_convert(d_lat, d_lat16, sites, bits, true, stream));
        lat16_bits = bits;
        narrow_valid = true;
        return CPM_OK;
    }

    // Centroids c = anchor + csum / area of the current lattice (settled
    // tables), into out [(n+1)*3]; dead cells get 0.
    int centroids_now(std::vector<double>& out, std::vector<int32_t>& areas) {
        const size_t rows = (size_t)n + 1;
        std::vector<int64_t> tab(rows), csum(rows * 3);
        std::vector<int32_t> anchor(rows * 3);
        if (int rc = ensure_wide()) return rc;
        HIP_TRY(hipMemsetAsync(d_csum, 0, rows * 3 * 8, stream));
        HIP_TRY(gcpm::launch_csum(d, d_lat, lgL, d_anchor, d_csum, stream));
        HIP_TRY(hipMemcpyAsync(tab.data(), d_tab[0], rows * 8, hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipMemcpyAsync(csum.data(), d_csum, rows * 3 * 8, hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipMemcpyAsync(anchor.data(), d_anchor, rows * 3 * 4, hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        out.assign(rows * 3, 0.0);
        areas.assign(rows, 0);
        for (size_t c = 0; c < rows; c++) {
            const int32_t area = (int32_t)(uint32_t)tab[c];
            areas[c] = area;
            for (int a = 0; a < 3; a++)
                out[3 * c + a] = (a < d && area > 0)
                                     ? (double)anchor[3 * c + a] + (double)csum[3 * c + a] / (double)area
                                     : 0.0;
        }
        return CPM_OK;
    }

    int upload_pref() {
        std::vector<int32_t> q(pref.size());
        for (size_t i = 0; i < q.size(); i++) q[i] = (int32_t)std::llround(pref[i] * 65536.0);
        HIP_TRY(hipMemcpyAsync(d_pref, q.data(), q.size() * 4, hipMemcpyHostToDevice, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        pref_dirty = false;
        return CPM_OK;
    }

    // Fill the history ring with the current centroids (first persistence run
    // after a push).
    int init_ring() {
        std::vector<double> c;
        std::vector<int32_t> areas;
        if (int rc = centroids_now(c, areas)) return rc;
        ring.resize((size_t)history * c.size());
        for (int i = 0; i < history; i++) std::copy(c.begin(), c.end(), ring.begin() + (size_t)i * c.size());
        ring_active = 0;
        ring_valid = true;
        return CPM_OK;
    }

    // Preferred-direction update after a run chunk (kernels.cu:158-212,
    // cpm.cu:487-514; displacement wrapped before its length is taken).
    int update_pref_dirs() {
        std::vector<double> now;
        std::vector<int32_t> areas;
        if (int rc = centroids_now(now, areas)) return rc;
        const size_t m = now.size();
        std::copy(now.begin(), now.end(), ring.begin() + (size_t)ring_active * m);
        ring_active = (ring_active + 1) % history;
        const double* prev = ring.data() + (size_t)ring_active * m;
        for (int c = 1; c <= n; c++) {
            if (areas[c] <= 0) continue;
            double u[3];
            for (int a = 0; a < 3; a++) u[a] = wrap_half(now[3 * c + a] - prev[3 * c + a], L);
            const double len = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
            if (len == 0.0) continue;
            const double D = st.pers_diff[ctype[c]];
            double np[3];
            for (int a = 0; a < 3; a++) np[a] = u[a] / len * (1.0 - D) + pref[3 * c + a] * D;
            const double nl = std::sqrt(np[0] * np[0] + np[1] * np[1] + np[2] * np[2]);
            if (nl == 0.0) continue;
            for (int a = 0; a < 3; a++) pref[3 * c + a] = np[a] / nl;
        }
        return upload_pref();
    }

    bool act_on() const {
        for (int t = 0; t < CPM_NTYPES; t++)
            if (st.lam_act[t] != 0.0) return true;
        return false;
    }

    bool chem_on() const {
        for (int t = 0; t < CPM_NTYPES; t++)
            if (st.chem[t] != 0.0) return true;
        return false;
    }

    // Smallest dyadic scale 2^k making every lambda integral (k <= 8); with a
    // chemotaxis term 2^8, which resolves it to 1/256 energy units.
    int energy_scale() const {
        for (int k = chem_on() || act_on() || pers_on() ? kMaxScaleLog2 : 0; k <= kMaxScaleLog2; k++) {
            const double s = (double)(1 << k);
            bool ok = true;
            for (int t = 0; t < CPM_NTYPES && ok; t++) {
                for (double v : {st.lam_area[t] * s, st.lam_perim[t] * s})
                    if (v != std::floor(v) || std::fabs(v) >= 2147483648.0) ok = false;
                for (int o = 0; o < CPM_NTYPES; o++) {
                    const double v = st.J[t][o] * s;
                    if (v != std::floor(v) || std::fabs(v) >= 1073741824.0) ok = false;
                }
            }
            if (ok) return 1 << k;
        }
        return -1;
    }

    // Launch form for a run with checkerboard edge c: one launch per phase
    // unless cpm_set_launch asks for a persistent form (and tile edge).
    int plan(int c, Plan* p) {
        if (want_form < Plan::kGlobal || want_form >= Plan::kSmall) {
            p->mode = Plan::kPhase;
            return CPM_OK;
        }
        const bool want_lds = want_form == Plan::kLds2d;
        if (d == 2 && want_lds) {
            int t = std::max(2 * c, std::min(64, L / 16));
            if (want_tile > 0) t = want_tile;
            if (t > L) t = L;
            if (is_pow2(t) && t >= 2 * c) {
                const int lt = ilog2(t);
                const int blocks = 1 << (2 * (lgL - lt));
                const size_t bytes = (size_t)(t + 2) * (t + 2) * 4;
                int maxb = 0;
                HIP_TRY(gcpm::persist_max_blocks(2, true, bytes, &maxb));
                if (blocks <= maxb) {
                    p->mode = Plan::kLds2d;
                    p->blocks = blocks;
                    p->ltile = lt;
                    p->lds_bytes = bytes;
                    return CPM_OK;
                }
            }
            return fail(CPM_EINVAL, "launch form 2: the LDS tiles do not fit co-resident");
        }
        if (d == 3) {
            // The persistent global-memory form is switched off in 3D: at
            // 256^3 (2048 co-resident blocks polling one barrier word) a run
            // did not finish within its time limit.  3D runs one launch per
            // colour phase.
            return fail(CPM_ENOTIMPL, "the persistent launch forms are disabled in 3D");
        }
        int maxb = 0;
        HIP_TRY(gcpm::persist_max_blocks(d, false, 0, &maxb));
        const long nsub = 1L << ((lgL - ilog2(c) - 1) * d);
        const long work = std::max(nsub, (long)n + 1);
        p->mode = Plan::kGlobal;
        p->blocks = (int)std::max(1L, std::min((long)maxb, (work + gcpm::kBlock - 1) / gcpm::kBlock));
        return CPM_OK;
    }

    int upload_params() {
        if (!params_dirty) return CPM_OK;
        const int sc = energy_scale();
        if (sc < 0)
            return fail(CPM_EINVAL,
                        "lambda_area / lambda_perimeter / adhesion must be multiples of 2^-8 "
                        "(dyadic fixed-point energy)");
        // Compact block of the nt types present (layout in cpm_device.h).
        gcpm::DevParams p{};
        const int n2 = nt * nt;
        for (int a = 0; a < nt; a++) {
            for (int b = 0; b < nt; b++) {
                p.w[a * nt + b] = (int32_t)(st.J[a][b] * sc);
                p.w[n2 + a * nt + b] = (int32_t)(st.J[a][b] * sc) + (int32_t)st.S[a][b] * sc;
            }
            p.w[2 * n2 + a] = (int32_t)(st.lam_area[a] * sc);
            p.w[2 * n2 + nt + a] = st.target_area[a];
            p.w[2 * n2 + 2 * nt + a] = (int32_t)(st.lam_perim[a] * sc);
            p.w[2 * n2 + 3 * nt + a] = st.target_perim[a];
        }
        for (int a = 0; a < CPM_NTYPES; a++)
            for (int kk = 0; kk < 4; kk++) {
                p.chem_m[a][kk] = kk ? (int64_t)std::llround(st.chem[a] * (double)sc * 256.0 /
                                                             std::sqrt((double)kk))
                                     : 0;
                p.pers_m[a][kk] = kk ? (int64_t)std::llround(st.lam_pers[a] * (double)sc * 256.0 /
                                                             std::sqrt((double)kk))
                                     : 0;
            }
        // Act: multipliers llround(lambda scale / actMax 2^16), the 2^x table
        // llround(2^(j/256) 2^30) and the log2 table llround(log2(v) 2^24).
        int amax = 0;
        for (int a = 0; a < CPM_NTYPES; a++) {
            p.act_max[a] = st.act_max[a];
            p.act_m[a] = st.act_max[a] >= 1 ? (int64_t)std::llround(st.lam_act[a] * (double)sc /
                                                                   (double)st.act_max[a] * 65536.0)
                                             : 0;
            amax = std::max(amax, st.act_max[a]);  // any type's G may be evaluated
        }
        if (act_on()) {
            std::vector<int64_t> tab(256 + (size_t)amax + 1);
            for (int j = 0; j < 256; j++) tab[j] = (int64_t)std::llround(std::exp2(j / 256.0) * 1073741824.0);
            tab[256] = 0;
            for (int v = 1; v <= amax; v++) tab[256 + v] = (int64_t)std::llround(std::log2((double)v) * 16777216.0);
            if ((int)tab.size() > act_tab_cap) {
                (void)hipFree(d_act_tab);
                d_act_tab = nullptr;
                HIP_TRY(hipMalloc(&d_act_tab, tab.size() * 8));
                act_tab_cap = (int)tab.size();
            }
            HIP_TRY(hipMemcpyAsync(d_act_tab, tab.data(), tab.size() * 8, hipMemcpyHostToDevice, stream));
            HIP_TRY(hipStreamSynchronize(stream));
        }
        if (nt <= gcpm::kSmallTypes) {
            constexpr int S = gcpm::kSmallTypes, S2 = S * S;
            for (int a = 0; a < nt; a++) {
                for (int b = 0; b < nt; b++) {
                    p.s4[a * S + b] = p.w[a * nt + b];
                    p.s4[S2 + a * S + b] = p.w[n2 + a * nt + b];
                }
                for (int f = 0; f < 4; f++) p.s4[2 * S2 + f * S + a] = p.w[2 * n2 + f * nt + a];
            }
        }
        // Acceptance thresholds thr[k] = floor(exp(-k / (scale T)) 2^32) for
        // a 32-bit draw; with the reference stream the draw u is MINSTD's and
        // the reference's r = u / (2^31 - 1) < exp(-E / T) is, exactly,
        // u < thr[k] = ceil(exp(-k / (scale T)) (2^31 - 1)) (kept while > 1).
        std::vector<uint32_t> thr(1, 0xFFFFFFFFu);
        if (st.temperature > 0.0) {
            for (int k = 1;; k++) {
                if (k >= kMaxThr)
                    return fail(CPM_EINVAL, "temperature * energy scale too large for the acceptance table");
                const double e = std::exp(-(double)k / ((double)sc * st.temperature));
                if (rng == 1) {
                    const double v = std::ceil(e * 2147483647.0);
                    if (v <= 1.0) break;
                    thr.push_back((uint32_t)v);
                } else {
                    const double v = std::floor(e * 4294967296.0);
                    if (v <= 0.0) break;
                    thr.push_back((uint32_t)v);
                }
            }
        }
        if ((int)thr.size() > thr_cap) {
            (void)hipFree(d_thr);
            d_thr = nullptr;
            HIP_TRY(hipMalloc(&d_thr, thr.size() * sizeof(uint32_t)));
            thr_cap = (int)thr.size();
        }
        HIP_TRY(hipMemcpyAsync(d_thr, thr.data(), thr.size() * sizeof(uint32_t), hipMemcpyHostToDevice, stream));
        HIP_TRY(hipMemcpyAsync(d_prm, &p, sizeof(p), hipMemcpyHostToDevice, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        nthr = (int)thr.size();
        scale = sc;
        params_dirty = false;
        return CPM_OK;
    }
};

// A fixed list of engines run together in one cooperative launch.
struct cpm_batch {
    std::vector<cpm_handle> sims;
    int d = 0;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    std::vector<hipEvent_t> ready;       // one per replica stream
    unsigned int* d_bars = nullptr;      // 2 words per replica, then the abandoned-wait flag
    double block_scale = 1.0;            // cpm_batch_set_blocks: blocks per unit of work
    int form = 0;                        // cpm_batch_set_form: -1 choose, 0 per phase, 1 persistent, 2 graph
    gcpm::ReplicaDesc* d_reps = nullptr;
    int* d_rep_of = nullptr;
    int rep_of_cap = 0;
    int last_blocks = 0;
    bool autotune = false;  // cpm_batch_set_autotune
    FormTuner tuner;
    KernelOpts opts;        // cpm_batch_set_kernel_options (small, compact, att)
    char last_kernel[96] = "";  // per-phase batch kernel of the last run
    // Graph replay (form 2): the instantiated graph of G MCS, its shape key
    // and the device MCS offset its kernels add.
    hipGraphExec_t gexec = nullptr;
    uint64_t gkey = 0;
    uint32_t* d_mbase = nullptr;
    int64_t graph_captures = 0, graph_replays = 0;
    int last_form = 0;  // the form the last run used
    // Kernel timing (cpm_batch_timing_begin/end): one event pair per launch.
    bool timing = false;
    std::vector<hipEvent_t> ev;
    size_t ev_used = 0;
    int64_t t_launches = 0, t_updates = 0;

    int next_events(hipEvent_t* e0, hipEvent_t* e1) {
        while (ev.size() < ev_used + 2) {
            hipEvent_t e;
            HIP_TRY(hipEventCreate(&e));
            ev.push_back(e);
        }
        *e0 = ev[ev_used];
        *e1 = ev[ev_used + 1];
        ev_used += 2;
        return CPM_OK;
    }

    ~cpm_batch() {
        if (stream) (void)hipStreamSynchronize(stream);
        (void)hipFree(d_bars);
        (void)hipFree(d_reps);
        (void)hipFree(d_rep_of);
        for (hipEvent_t e : ready) (void)hipEventDestroy(e);
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        if (gexec) (void)hipGraphExecDestroy(gexec);
        (void)hipFree(d_mbase);
        if (done) (void)hipEventDestroy(done);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

// After persistent launches: wait for them and read the abandoned-wait flag
// (the last of nwords barrier words).  A set flag means a block of the
// cooperative grid never became resident while the others waited; the
// launch ended early, so the run's results are invalid.  The words are reset
// for the next launch.
int check_barrier(unsigned int* words, int nwords, hipStream_t s) {
    unsigned int flag = 0;
    HIP_TRY(hipMemcpyAsync(&flag, words + nwords - 1, sizeof(flag), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (!flag) return CPM_OK;
    HIP_TRY(hipMemsetAsync(words, 0, sizeof(unsigned int) * nwords, s));
    HIP_TRY(hipStreamSynchronize(s));
    return fail(CPM_EHIP,
                "persistent launch: a grid barrier wait was abandoned (not every block of the "
                "cooperative grid was resident); the run's results are invalid");
}

int check_handle(cpm_handle h) {
    if (!h) return fail(CPM_EINVAL, "null handle");
    return CPM_OK;
}

int check_type(int t, const char* what) {
    if (t < 0 || t >= CPM_NTYPES)
        return fail(CPM_EINVAL, std::string(what) + " must be in [0, 20)");
    return CPM_OK;
}

}  // namespace

extern "C" {

const char* cpm_last_error(void) { return g_err.c_str(); }

const char* cpm_version(void) { return "gpucpm-mi355x 0.1 (gfx950)"; }

int cpm_create(int dimension, int dimensionality, int number_of_cells, int temperature,
               int has_field, int history, int has_act, cpm_handle* out) {
    if (!out) return fail(CPM_EINVAL, "null output handle");
    *out = nullptr;
    if (dimensionality != 2 && dimensionality != 3)
        return fail(CPM_EINVAL, "dimensionality must be 2 or 3");
    // getNeighbor masks with dimension-1 (random.cu:22-28): powers of two only.
    if (!is_pow2(dimension) || dimension < 2)
        return fail(CPM_EINVAL, "dimension must be a power of two >= 2");
    if ((long)ilog2(dimension) * dimensionality > 30)
        return fail(CPM_EINVAL, "lattice too large (at most 2^30 sites)");
    if (number_of_cells < 0 || number_of_cells >= (1 << 24))
        return fail(CPM_EINVAL, "number_of_cells must be in [0, 2^24)");
    if (history < 1) return fail(CPM_EINVAL, "history must be >= 1");

    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev < 1) return fail(CPM_EHIP, "no HIP device available");

    cpm_engine* g = new (std::nothrow) cpm_engine();
    if (!g) return fail(CPM_EHIP, "host allocation failed");
    g->L = dimension;
    g->d = dimensionality;
    g->n = number_of_cells;
    g->lgL = ilog2(dimension);
    g->sites = (size_t)1 << (g->lgL * g->d);
    g->st.temperature = (double)temperature;
    std::random_device rd;  // like cpm.cu:347-350, but once per engine
    g->seed = ((uint64_t)rd() << 32) ^ rd();
    g->history = history;
    g->pref.assign(((size_t)number_of_cells + 1) * 3, 0.0);
    g->ctype.assign((size_t)number_of_cells + 1, 0);
    g->init_pref_dirs();

    auto bail = [&](hipError_t err, const char* what) {
        int rc = hip_fail(err, what);
        delete g;
        return rc;
    };
    const size_t rows = (size_t)g->n + 1;
    hipError_t err;
    if ((err = hipStreamCreate(&g->stream)) != hipSuccess) return bail(err, "hipStreamCreate");
    if ((err = hipEventCreateWithFlags(&g->ev_ready, hipEventDisableTiming)) != hipSuccess)
        return bail(err, "hipEventCreate");
    if ((err = hipHostMalloc((void**)&g->h_lat, g->sites * 4, hipHostMallocDefault)) != hipSuccess)
        return bail(err, "hipHostMalloc(lattice)");
    std::memset(g->h_lat, 0, g->sites * 4);
    if ((err = hipMalloc(&g->d_lat, g->sites * 4)) != hipSuccess) return bail(err, "hipMalloc(lattice)");
    for (int i = 0; i < 2; i++)
        if ((err = hipMalloc(&g->d_tab[i], rows * 8)) != hipSuccess) return bail(err, "hipMalloc(tables)");
    for (int i = 0; i < 2; i++)
        if ((err = hipMalloc(&g->d_del[i], rows * 8)) != hipSuccess) return bail(err, "hipMalloc(deltas)");
    if ((err = hipMalloc(&g->d_anchor, rows * 3 * 4)) != hipSuccess) return bail(err, "hipMalloc(anchor)");
    if ((err = hipMalloc(&g->d_csum, rows * 3 * 8)) != hipSuccess) return bail(err, "hipMalloc(csum)");
    if ((err = hipMalloc(&g->d_first, rows * 4)) != hipSuccess) return bail(err, "hipMalloc(first)");
    if ((err = hipMalloc(&g->d_prm, sizeof(gcpm::DevParams))) != hipSuccess) return bail(err, "hipMalloc(params)");
    if ((err = hipMalloc(&g->d_bar, 3 * sizeof(unsigned int))) != hipSuccess) return bail(err, "hipMalloc(barrier)");
    if ((err = hipMemset(g->d_bar, 0, 3 * sizeof(unsigned int))) != hipSuccess) return bail(err, "hipMemset(barrier)");
    if ((err = hipMalloc(&g->d_pref, rows * 3 * 4)) != hipSuccess) return bail(err, "hipMalloc(pref)");
    if ((err = hipMalloc(&g->d_ctype, rows * 4)) != hipSuccess) return bail(err, "hipMalloc(ctype)");
    if (has_act) {  // cpm.cu:192-193,204-205
        g->has_act = true;
        if ((err = hipHostMalloc((void**)&g->h_act, g->sites * 4, hipHostMallocDefault)) != hipSuccess)
            return bail(err, "hipHostMalloc(act)");
        std::memset(g->h_act, 0, g->sites * 4);
        if ((err = hipMalloc(&g->d_act, g->sites * 4)) != hipSuccess) return bail(err, "hipMalloc(act)");
        if ((err = hipMemset(g->d_act, 0, g->sites * 4)) != hipSuccess) return bail(err, "hipMemset(act)");
    }
    if (has_field) {
        g->has_field = true;
        g->h_field.assign(g->sites * g->d, 0.0);
        g->h_field_pushed.assign(g->sites * g->d, 0.0);
        if ((err = hipMalloc(&g->d_field, g->sites * g->d * 4)) != hipSuccess) return bail(err, "hipMalloc(field)");
        if ((err = hipMemset(g->d_field, 0, g->sites * g->d * 4)) != hipSuccess) return bail(err, "hipMemset(field)");
    }
    *out = g;
    return CPM_OK;
}

int cpm_destroy(cpm_handle h) {
    delete h;
    return CPM_OK;
}

int cpm_set_temperature(cpm_handle h, double temperature) {
    if (int rc = check_handle(h)) return rc;
    if (!(temperature >= 0.0) || !std::isfinite(temperature))
        return fail(CPM_EINVAL, "temperature must be finite and >= 0");
    h->st.temperature = temperature;
    h->params_dirty = true;
    return CPM_OK;
}

int cpm_set_area(cpm_handle h, int type, int target, double lambda) {
    if (int rc = check_handle(h)) return rc;
    if (int rc = check_type(type, "cell_type")) return rc;
    h->st.lam_area[type] = lambda;
    h->st.target_area[type] = target;
    h->params_dirty = true;
    return CPM_OK;
}

int cpm_set_perimeter(cpm_handle h, int type, int target, double lambda) {
    if (int rc = check_handle(h)) return rc;
    if (int rc = check_type(type, "cell_type")) return rc;
    h->st.lam_perim[type] = lambda;
    h->st.target_perim[type] = target;
    h->params_dirty = true;
    return CPM_OK;
}

int cpm_set_adhesion(cpm_handle h, int type, int other_type, int adhesion, int set_both) {
    if (int rc = check_handle(h)) return rc;
    if (int rc = check_type(type, "cell_type")) return rc;
    if (int rc = check_type(other_type, "other_cell_type")) return rc;
    h->st.J[type][other_type] = (int16_t)adhesion;  // short, settings.h:20
    if (set_both) h->st.J[other_type][type] = (int16_t)adhesion;
    h->params_dirty = true;
    return CPM_OK;
}

int cpm_set_adhesion_real(cpm_handle h, int type, int other_type, double adhesion, int set_both) {
    if (int rc = check_handle(h)) return rc;
    if (int rc = check_type(type, "cell_type")) return rc;
    if (int rc = check_type(other_type, "other_cell_type")) return rc;
    if (!std::isfinite(adhesion) || std::fabs(adhesion) > 32767.0)
        return fail(CPM_EINVAL, "adhesion must be finite with |J| <= 32767");
    h->st.J[type][other_type] = adhesion;
    if (set_both) h->st.J[other_type][type] = adhesion;
    h->params_dirty = true;
    return CPM_OK;
}

int cpm_set_stickiness(cpm_handle h, int type, int other_type, int stickiness) {
    if (int rc = check_handle(h)) return rc;
    if (int rc = check_type(type, "cell_type")) return rc;
    if (int rc = check_type(other_type, "other_cell_type")) return rc;
    h->st.S[type][other_type] = (int16_t)stickiness;
    h->params_dirty = true;
    return CPM_OK;
}

int cpm_set_fixed(cpm_handle h, int type, int fixed) {
    if (int rc = check_handle(h)) return rc;
    if (int rc = check_type(type, "cell_type")) return rc;
    h->st.fixed[type] = fixed;
    return CPM_OK;
}

int cpm_set_chemotaxis(cpm_handle h, int type, double lambda) {
    if (int rc = check_handle(h)) return rc;
    if (int rc = check_type(type, "cell_type")) return rc;
    if (!std::isfinite(lambda) || std::fabs(lambda) > 8192.0)
        return fail(CPM_EINVAL, "lambda_chemotaxis must be finite with |lambda| <= 8192");
    h->st.chem[type] = lambda;
    h->params_dirty = true;
    return CPM_OK;
}

int cpm_set_act(cpm_handle h, int type, int max_act, double lambda) {
    if (int rc = check_handle(h)) return rc;
    if (int rc = check_type(type, "cell_type")) return rc;
    if (!std::isfinite(lambda)) return fail(CPM_EINVAL, "lambda_act must be finite");
    if (max_act < 0 || max_act > 65535) return fail(CPM_EINVAL, "max_act must be in [0, 65535]");
    h->st.lam_act[type] = lambda;
    h->st.act_max[type] = max_act;
    h->params_dirty = true;
    return CPM_OK;
}

int cpm_host_act(cpm_handle h, uint32_t** out) {
    if (int rc = check_handle(h)) return rc;
    if (!out) return fail(CPM_EINVAL, "null output");
    if (!h->has_act) return fail(CPM_ESTATE, "this engine was created with has_act = 0");
    *out = h->h_act;
    return CPM_OK;
}

int cpm_set_persistence(cpm_handle h, int type, int time, double diffusion, double lambda) {
    if (int rc = check_handle(h)) return rc;
    if (int rc = check_type(type, "cell_type")) return rc;
    if (!std::isfinite(lambda) || std::fabs(lambda) > 8192.0)
        return fail(CPM_EINVAL, "lambda_persistence must be finite with |lambda| <= 8192");
    if (!std::isfinite(diffusion)) return fail(CPM_EINVAL, "persistence_diffusion must be finite");
    h->st.pers_time[type] = time;
    h->st.pers_diff[type] = diffusion;
    h->st.lam_pers[type] = lambda;
    h->params_dirty = true;
    return CPM_OK;
}

int cpm_get_pref_dirs(cpm_handle h, double* out) {
    if (int rc = check_handle(h)) return rc;
    if (!out) return fail(CPM_EINVAL, "null output");
    std::memcpy(out, h->pref.data(), h->pref.size() * sizeof(double));
    return CPM_OK;
}

int cpm_set_pref_dirs(cpm_handle h, const double* in) {
    if (int rc = check_handle(h)) return rc;
    if (!in) return fail(CPM_EINVAL, "null input");
    for (size_t i = 0; i < h->pref.size(); i++)
        if (!(std::fabs(in[i]) < 32768.0)) return fail(CPM_EINVAL, "preferred directions must satisfy |p| < 2^15");
    std::memcpy(h->pref.data(), in, h->pref.size() * sizeof(double));
    h->pref_dirty = true;
    return CPM_OK;
}

int cpm_host_field(cpm_handle h, double** out) {
    if (int rc = check_handle(h)) return rc;
    if (!out) return fail(CPM_EINVAL, "null output");
    if (!h->has_field) return fail(CPM_ESTATE, "this engine was created with has_field = 0");
    *out = h->h_field.data();
    return CPM_OK;
}

int cpm_add_cell(cpm_handle h, int type, int x, int y, int z) {
    if (int rc = check_handle(h)) return rc;
    if (int rc = check_type(type, "cell type")) return rc;
    if (h->active >= h->n)
        return fail(CPM_EINVAL, "add_cell: more cells than number_of_cells");
    const int L = h->L;
    if (x < 0 || x >= L || y < 0 || y >= L) return fail(CPM_EINVAL, "add_cell: position out of range");
    size_t idx;
    if (h->d == 3) {
        if (z < 0 || z >= L) return fail(CPM_EINVAL, "add_cell: z out of range");
        idx = ((size_t)z * L + y) * L + x;
    } else {
        idx = (size_t)y * L + x;
    }
    h->active++;
    h->h_lat[idx] = (uint32_t)h->active | ((uint32_t)type << 24);
    return CPM_OK;
}

int cpm_host_lattice(cpm_handle h, uint32_t** out) {
    if (int rc = check_handle(h)) return rc;
    if (!out) return fail(CPM_EINVAL, "null output");
    *out = h->h_lat;
    return CPM_OK;
}

int cpm_active_cells(cpm_handle h, int* out) {
    if (int rc = check_handle(h)) return rc;
    if (!out) return fail(CPM_EINVAL, "null output");
    *out = h->active;
    return CPM_OK;
}

int cpm_push(cpm_handle h) {
    if (int rc = check_handle(h)) return rc;
    uint32_t tmax = 0;
    for (size_t i = 0; i < h->sites; i++) {
        const uint32_t v = h->h_lat[i];
        if ((v & gcpm::kSpinMask) > (uint32_t)h->n)
            return fail(CPM_EINVAL, "lattice holds a cell number above number_of_cells");
        if ((v >> 24) >= (uint32_t)CPM_NTYPES)
            return fail(CPM_EINVAL, "lattice holds a cell type >= 20");
        tmax = std::max(tmax, v >> 24);
    }
    if ((int)tmax + 1 != h->nt) {
        h->nt = (int)tmax + 1;  // flips only copy existing values: no new types appear
        h->params_dirty = true;
    }
    for (size_t i = 0; i < h->sites; i++) {  // a cell keeps its type
        const uint32_t v = h->h_lat[i];
        if (v & gcpm::kSpinMask) h->ctype[v & gcpm::kSpinMask] = (int32_t)(v >> 24);
    }
    h->ring_valid = false;
    // The field goes with the lattice (cpm.cu:542-543), quantized to 2^-16.
    std::vector<int32_t> fq;
    if (h->has_field) {
        fq.resize(h->h_field.size());
        for (size_t i = 0; i < fq.size(); i++) {
            const double f = h->h_field[i];
            if (!(std::fabs(f) < 32768.0)) return fail(CPM_EINVAL, "field values must satisfy |f| < 2^15");
            fq[i] = (int32_t)std::llround(f * 65536.0);
        }
    }
    const size_t rows = (size_t)h->n + 1;
    hipStream_t s = h->stream;
    if (h->has_field) {
        HIP_TRY(hipMemcpyAsync(h->d_field, fq.data(), fq.size() * 4, hipMemcpyHostToDevice, s));
        h->h_field_pushed = h->h_field;
    }
    if (h->has_act)  // cpm.cu:548
        HIP_TRY(hipMemcpyAsync(h->d_act, h->h_act, h->sites * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(h->d_lat, h->h_lat, h->sites * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(h->d_ctype, h->ctype.data(), rows * 4, hipMemcpyHostToDevice, s));
    h->wide_valid = true;
    h->narrow_valid = false;
    h->res_auto_off = false;  // a new lattice: the resident form's launches start at full length
    h->res_small_tiles = false;
    h->res_mcs = 0;
    h->res_streak = 0;
    HIP_TRY(hipMemsetAsync(h->d_tab[0], 0, rows * 8, s));
    for (int i = 0; i < 2; i++) HIP_TRY(hipMemsetAsync(h->d_del[i], 0, rows * 8, s));
    HIP_TRY(hipMemsetAsync(h->d_first, 0xFF, rows * 4, s));
    HIP_TRY(gcpm::launch_tally(h->d, h->d_lat, h->lgL, h->n, h->d_tab[0], h->d_first, h->d_anchor, s));
    HIP_TRY(hipMemcpyAsync(h->d_tab[1], h->d_tab[0], rows * 8, hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    h->pushed = true;
    return CPM_OK;
}

namespace {

// Active sub-block offsets of phase pc of a round: colour (cbX outermost,
// kernels2d.cu:179-180) then group, offset = group K + colour per axis.
// With block slices (k.lg > 0, shared roles: S = 1) phase pc of a round is
// slice pc / K^d (phase_slice), colour pc mod K^d.
void phase_offsets(int d, const gcpm::SweepConsts& k, int pc, int* ox, int* oy, int* oz) {
    const int K = k.K, S = k.S;
    if (k.lg) pc %= d == 3 ? K * K * K : K * K;
    const int ngrp = d == 3 ? S * S * S : S * S;
    const int col = pc / ngrp, grp = pc % ngrp;
    const int cx = d == 3 ? col / (K * K) : col / K;
    const int cy = d == 3 ? (col / K) % K : col % K;
    const int cz = d == 3 ? col % K : 0;
    const int gx = d == 3 ? grp / (S * S) : grp / S;
    const int gy = d == 3 ? (grp / S) % S : grp % S;
    const int gz = d == 3 ? grp % S : 0;
    *ox = gx * K + cx;
    *oy = gy * K + cy;
    *oz = gz * K + cz;
}

void phase_slice(int d, const gcpm::SweepConsts& k, int pc, int* gx, int* gy, int* gz) {
    *gx = *gy = *gz = 0;
    if (!k.lg) return;
    const int K = k.K, G = 1 << k.lg;
    const int gs = pc / (d == 3 ? K * K * K : K * K);  // x outermost (kernels2d.cu:92-93)
    *gx = d == 3 ? gs / (G * G) : gs / G;
    *gy = d == 3 ? (gs / G) % G : gs % G;
    *gz = d == 3 ? gs % G : 0;
}

// Validate run arguments against engine h and fill the sweep constants
// (cpm.cu:291-405 semantics; see cpm_run).
int prepare_run(cpm_handle h, const cpm_run_args* a, gcpm::SweepConsts* kout, int* lh_out) {
    if (int rc = check_handle(h)) return rc;
    if (!a) return fail(CPM_EINVAL, "null run arguments");
    if (!h->pushed) return fail(CPM_ESTATE, "run before push_to_gpu");
    if (a->threads_per_block < 1 || a->positions_per_thread < 1 || a->updates_per_barrier < 1)
        return fail(CPM_EINVAL, "threads_per_block, positions_per_thread and updates_per_barrier must be >= 1");
    const int c = a->positions_per_checkerboard;
    if (!is_pow2(c) || h->L % (2 * c))
        return fail(CPM_EINVAL, "positions_per_checkerboard must be a power of two with 2*c dividing dimension");
    // Schedule (kernels2d.cu:72-106,179-243): shared roles use regions of 2c
    // sites and K = 2 colours per axis; non-shared roles (the reference's
    // default, shared = 0) regions of positions_per_thread sites, K = ppt / c,
    // with the regions of a colour in S^d groups so that concurrent attempts
    // stay >= 3 sites apart (S ppt - c >= 3, S <= L / ppt).
    const bool shared = a->shared != 0;
    const int R = shared ? 2 * c : a->positions_per_thread;
    if (!shared && (R < c || R % c || h->L % R || !is_pow2(R)))
        return fail(CPM_EINVAL,
                    "with shared = 0, positions_per_thread must be a multiple of "
                    "positions_per_checkerboard dividing dimension");
    const int K = R / c;
    int S = 1;
    if (!shared) {
        while (S * R - c < 3) S *= 2;
        if (S > h->L / R) S = h->L / R;
    }
    // The reference stream with shared roles and positions_per_thread > 2c
    // keeps the reference's thread layout (kernels2d.cu:72-86,92-106,136-146;
    // cpm.cu:363): L / ppt threads per axis in blocks of threads_per_block,
    // each owning ppt / 2c regions of 2c sites per axis, visited one block
    // slice at a time; its stream is the thread's.
    int lg = 0, ld = 0;
    const bool gblocks = h->rng == 1 && shared && a->positions_per_thread > R;
    if (gblocks) {
        const int ppt = a->positions_per_thread, tpb = a->threads_per_block;
        if (ppt % R || h->L % ppt || (h->L / ppt) % tpb)
            return fail(CPM_EINVAL,
                        "the reference stream with shared = 1 and positions_per_thread > "
                        "2 * positions_per_checkerboard needs 2c to divide positions_per_thread and "
                        "threads_per_block * positions_per_thread to divide dimension (the reference's grid)");
        lg = ilog2(ppt / R);
        ld = ilog2(tpb);
    }
    const long cd = (long)c * c * (h->d == 3 ? c : 1);
    const int u = a->updates_per_checkerboard_switch;
    if (u < 1 || u > cd)
        return fail(CPM_EINVAL, "updates_per_checkerboard_switch must be in [1, c^d]");
    if (a->inner_iterations < 1) return fail(CPM_EINVAL, "inner_iterations must be >= 1");
    if (a->iterations < 0) return fail(CPM_EINVAL, "iterations must be >= 0");
    if (int rc = h->upload_params()) return rc;
    gcpm::SweepConsts k{};
    k.lat = h->d_lat;
    k.lat16 = nullptr;  // set by the caller when it runs on the 16-bit lattice
    k.counters = h->counters_on ? h->d_counters : nullptr;
    k.idbits = 24;
    k.idmask = gcpm::kSpinMask;
    for (int i = 0; i < 2; i++) k.tab[i] = h->d_tab[i];
    for (int i = 0; i < 2; i++) k.del[i] = h->d_del[i];
    k.prm = h->d_prm;
    k.thr = h->d_thr;
    k.nthr = h->nthr;
    k.ncells = h->n;
    k.lgL = h->lgL;
    k.lc = ilog2(c);
    k.u = u;
    k.rounds = (int)(cd / u);  // kernels2d.cu:95
    k.k0 = (uint32_t)h->seed;
    k.k1 = (uint32_t)(h->seed >> 32);
    k.nt = h->nt;
    k.chem = h->chem_on() ? 1 : 0;
    if (k.chem && !h->has_field)
        return fail(CPM_EINVAL, "lambda_chemotaxis is set but the engine has no field (has_field = 0)");
    k.field = h->d_field;
    k.act = h->act_on() ? 1 : 0;
    if (k.act) {
        if (!h->has_act)
            return fail(CPM_EINVAL, "lambda_act is set but the engine has no act lattice (has_act = 0)");
        for (int t = 0; t < CPM_NTYPES; t++) {
            if (h->st.lam_act[t] == 0.0) continue;
            if (h->st.act_max[t] < 1) return fail(CPM_EINVAL, "max_act must be >= 1 where lambda_act != 0");
            if (std::fabs(h->st.lam_act[t] * h->scale / h->st.act_max[t]) >= 16384.0)
                return fail(CPM_EINVAL, "|lambda_act / max_act| too large (at most 63)");
        }
    }
    // Persistence acts in 3D only, as in the reference (kernels3d.cu:327-344;
    // the prefDir update runs for dimensionality 3, cpm.cu:487).
    k.pers = h->d == 3 && h->pers_on() ? 1 : 0;
    k.pref = h->d_pref;
    if (k.pers && h->pref_dirty)
        if (int rc = h->upload_pref()) return rc;
    k.act_lat = h->d_act;
    k.act_tab = h->d_act_tab;
    k.act_ln2q = (int64_t)std::llround(std::log(2.0) * 1073741824.0);
    k.fixed_mask = 0;
    for (int t = 0; t < CPM_NTYPES; t++)
        if (h->st.fixed[t]) k.fixed_mask |= 1u << t;
    k.roles = shared ? 0 : 1;
    k.lstride = ilog2((long)S * K);
    k.K = K;
    k.S = S;
    k.phases_per_round = 1;
    for (int i = 0; i < h->d; i++) k.phases_per_round *= K * S << lg;  // (block slices x colours)
    k.refstream = 0;
    k.rstate = nullptr;
    k.lk = ilog2(K);
    k.lg = lg;
    k.ld = ld;
    k.lr = h->lgL - ilog2(R) - lg;
    *kout = k;
    *lh_out = h->lgL - k.lc - k.lstride - lg;
    return CPM_OK;
}

}  // namespace

int cpm_run(cpm_handle h, const cpm_run_args* a) {
    gcpm::SweepConsts k{};
    int lh = 0;
    if (int rc = prepare_run(h, a, &k, &lh)) return rc;
    const int c = a->positions_per_checkerboard;
    const int u = a->updates_per_checkerboard_switch;
    // cpm.cu:405: iterations / inner_iterations launches of inner_iterations MCS.
    const long chunks = a->iterations / a->inner_iterations;
    const long total = chunks * a->inner_iterations;
    if (total == 0) return CPM_OK;

    Plan plan;
    if (a->cell_sync) {
        plan.mode = Plan::kPhase;  // two launches per phase: claim, commit
        if (!h->d_claims) {
            HIP_TRY(hipMalloc(&h->d_claims, sizeof(uint64_t) * ((size_t)h->n + 1)));
            HIP_TRY(hipMemsetAsync(h->d_claims, 0, sizeof(uint64_t) * ((size_t)h->n + 1), h->stream));
        }
    } else if (k.roles) {
        plan.mode = Plan::kPhase;  // the persistent forms implement shared roles only
    } else if (int rc = h->plan(c, &plan)) {
        return rc;
    }
    // Act with shared roles and c = 1: an attempt's act term reads the
    // source's neighbourhood, two sites from its target, where other targets
    // of the phase lie.  Each phase is then a decide launch (every attempt
    // against the phase-start lattice, decisions recorded) and an apply
    // launch (the accepted ones committed; a target's own neighbourhood
    // holds no other target, so the deltas are those decided).  With
    // cell_sync the decide launch also claims and the apply launch checks.
    const bool split = k.act && !k.roles && c == 1;
    if (split) {
        plan = Plan{};
        plan.mode = Plan::kPhase;
        const size_t nsub = (size_t)1 << (lh * h->d);
        if (nsub > h->decide_cap) {
            (void)hipFree(h->d_decide);
            h->d_decide = nullptr;
            HIP_TRY(hipMalloc(&h->d_decide, nsub));
            h->decide_cap = nsub;
        }
    }
    if (h->rng == 1) {
        // The reference stream: one launch per colour phase (each reference
        // thread's state is carried between phases in d_rstate).
        plan = Plan{};
        plan.mode = Plan::kPhase;
        const size_t nreg = (size_t)1 << (h->d * k.lr);
        if (nreg > h->rstate_cap) {
            (void)hipFree(h->d_rstate);
            h->d_rstate = nullptr;
            HIP_TRY(hipMalloc(&h->d_rstate, nreg * sizeof(int32_t)));
            h->rstate_cap = nreg;
        }
        k.refstream = 1;
        k.rstate = h->d_rstate;
    }
    FormTuner::Entry* run_tune = nullptr;  // run-level autotuning of the launch form
    int run_ti = -1;
    hipEvent_t run_e0 = nullptr, run_e1 = nullptr;
    auto finish_run_tune = [&]() -> int {
        if (!run_tune) return CPM_OK;
        HIP_TRY(hipEventRecord(run_e1, h->stream));
        FormTuner::Pending pd{run_tune, run_ti, run_e0, run_e1};
        pd.scale = (double)total;
        h->tuner.pending.push_back(pd);
        HIP_TRY(h->tuner.settle());
        return CPM_OK;
    };
    // The single-workgroup form (lattice, tables and phases in one
    // workgroup's LDS): forced by launch form 3, chosen automatically when a
    // phase has at most kSmallAutoAttempts attempts (a whole chunk is then
    // cheaper on one CU than a launch per phase); not with cell_sync, the
    // reference stream or the act term.
    {
        const int bits_s = h->narrow_bits();
        const size_t bytes = gcpm::small_lds_bytes(h->d, h->lgL, h->n, bits_s != 0);
        const bool fits = bytes <= kSmallLdsBudget;
        const bool ok = !a->cell_sync && !k.refstream && !k.act && fits;
        const long attempts = ((long)1 << (lh * h->d)) * u;
        if (h->want_form == Plan::kSmall && !ok)
            return fail(CPM_EINVAL,
                        "launch form 3 (single workgroup) needs a lattice whose sites and tables fit in "
                        "one workgroup's LDS, and no cell_sync, reference stream or act term");
        bool use = ok && (h->want_form == Plan::kSmall || (h->want_form == -1 && attempts <= kSmallAutoAttempts));
        // Autotuner: where the form is eligible, whole runs alternate
        // between it and per-phase launches (each run timed per MCS) until
        // both have their samples; then the faster is kept for the workload.
        if (ok && h->autotune && h->want_form == -1 && !h->timing) {
            const uint64_t key = (uint64_t)3 << 62 | (uint64_t)h->d | (uint64_t)h->lgL << 2 |
                                 (uint64_t)k.lc << 8 | (uint64_t)u << 12 | (uint64_t)k.roles << 24 |
                                 (uint64_t)h->nt << 25 | (uint64_t)k.chem << 30 | (uint64_t)k.pers << 31 |
                                 (uint64_t)k.S << 32 | (uint64_t)(h->narrow_bits() != 0) << 40 |
                                 (uint64_t)a->inner_iterations << 41;
            run_tune = &h->tuner.entry(key, {0, 1});  // 0 per phase, 1 single workgroup
            if (run_tune->best >= 0) {
                use = run_tune->best == 1;
                run_tune = nullptr;
            } else {
                run_ti = h->tuner.pick(*run_tune);
                if (run_ti < 0) {
                    run_tune = nullptr;  // (all samples pending)
                } else {
                    use = run_tune->forms[run_ti] == 1;
                    HIP_TRY(h->tuner.events(&run_e0, &run_e1));
                    HIP_TRY(hipEventRecord(run_e0, h->stream));
                }
            }
        }
        if (use) {
            plan = Plan{};
            plan.mode = Plan::kSmall;
            plan.blocks = 1;
        }
    }
    // The resident form (2D, u = 1): launch form 4, and the automatic choice
    // for such runs on lattices of 256^2 sites and more whose phases are too
    // large for one workgroup (unless the autotuner decides, or a guard
    // failure that shorter launches cannot cure switched it off for the
    // lattice pushed).
    int res_lte = 0, res_var = 0;
    const bool res_ok = h->d == 2 && u == 1 && !a->cell_sync && !k.refstream && !k.act && h->L >= 128;
    const bool res_auto = h->want_form == -1 && plan.mode == Plan::kPhase && res_ok && h->L >= 256 &&
                          !h->autotune && !h->res_auto_off;
    if (h->want_form == Plan::kResident || res_auto) {
        if (!res_ok)
            return fail(CPM_EINVAL,
                        "launch form 4 (resident tiles) needs 2D, updates_per_checkerboard_switch = 1, L >= 128, "
                        "and no cell_sync, reference stream or act term");
        // tiles of 64 sites from 256^2 up (a phase takes the same time on
        // 16 tiles as on 256: each tile's chain of draws, prepare, commit
        // and exchange sets it, and 64^2 tiles run 1024 threads; measured
        // 7-17% faster than 32^2 tiles at 256^2 to 1024^2), 32 at 128^2:
        // at least four per axis.  set_launch(4, 32 or 64) picks the edge.
        res_lte = h->lgL >= 8 && !h->res_small_tiles ? 6 : 5;
        if (h->want_form == Plan::kResident && (h->want_tile == 32 || h->want_tile == 64) &&
            h->lgL - (h->want_tile == 64 ? 6 : 5) >= 2)
            res_lte = h->want_tile == 64 ? 6 : 5;
        // (the lean path: variant 0 for c = 1 with shared roles, 3 for the
        // other schedules)
        res_var = k.chem ? 2
                         : (k.nt <= gcpm::kSmallTypes && !k.fixed_mask && !h->counters_on &&
                            k.nthr <= gcpm::kResThrLean)
                               ? (k.lc == 0 && !k.roles ? 0 : 3)
                               : 1;
        int maxb = 0;
        HIP_TRY(gcpm::resident_max_blocks(h->narrow_bits() != 0, res_lte, res_var, h->n, &maxb));
        const int tiles = 1 << (2 * (h->lgL - res_lte));
        if (tiles > maxb && !res_auto) return fail(CPM_EINVAL, "launch form 4: the tiles do not fit co-resident");
        if (tiles <= maxb) {
            plan = Plan{};
            plan.mode = Plan::kResident;
            plan.blocks = tiles;
            plan.ltile = res_lte;
        }
    }
    h->last_plan = plan;
    const int64_t updates_per_mcs = (int64_t)k.rounds * k.phases_per_round * ((int64_t)1 << (lh * h->d)) * u;

    if (plan.mode == Plan::kResident) {
        const int bits = h->narrow_bits();
        gcpm::ResidentArgs ra{};
        ra.k = k;
        if (bits) {
            if (int rc = h->ensure_narrow(bits)) return rc;
            ra.k.lat16 = h->d_lat16;
            ra.k.idbits = (uint32_t)bits;
            ra.k.idmask = (1u << bits) - 1u;
        } else {
            if (int rc = h->ensure_wide()) return rc;
            h->narrow_valid = false;
        }
        int tiles = plan.blocks;
        const size_t rows = (size_t)h->n + 1;
        const size_t lat_bytes = h->sites * (bits ? 2 : 4);
        size_t nmail = 0;
        auto mail_for = [&]() -> int {
            nmail = (size_t)tiles * 16 * gcpm::res_mail_stride(1 << res_lte);
            if (nmail > h->res_mail_cap) {
                (void)hipFree(h->d_res_mail);
                h->d_res_mail = nullptr;
                h->res_mail_cap = 0;
                HIP_TRY(hipMalloc(&h->d_res_mail, sizeof(unsigned long long) * nmail));
                h->res_mail_cap = nmail;
            }
            return CPM_OK;
        };
        if (int rc = mail_for()) return rc;
        if (!h->d_res_flags) HIP_TRY(hipMalloc(&h->d_res_flags, sizeof(unsigned int) * (2 + rows)));
        if (!h->d_bbox) HIP_TRY(hipMalloc(&h->d_bbox, sizeof(int32_t) * 4 * rows));
        if (!h->d_backup_lat) HIP_TRY(hipMalloc(&h->d_backup_lat, h->sites * 4));
        if (!h->d_backup_tab) HIP_TRY(hipMalloc(&h->d_backup_tab, sizeof(int64_t) * rows));
        unsigned int* err = h->d_res_flags;
        unsigned int* bad = err + 1;
        ra.lte = res_lte;
        ra.lnt = h->lgL - res_lte;
        ra.gen = !(k.lc == 0 && !k.roles);
        ra.nmcs = a->inner_iterations;
        ra.holders = h->d_res_flags + 2;
        ra.err = err;
        ra.bad = bad;
        void* cur_lat = bits ? (void*)h->d_lat16 : (void*)h->d_lat;
        // Launches of up to K MCS (the results do not depend on how a run is
        // split).  K starts at the run's inner_iterations; a launch that
        // fails its guard (a cell met outside the tiles holding it, or wider
        // than two tiles) is redone with one launch per phase and halves K
        // for this engine; four launches in a row that hold double it again.
        const uint64_t base = h->mcs;
        const int64_t kmax = a->inner_iterations;
        int64_t K = h->res_mcs > 0 ? std::min<int64_t>(h->res_mcs, kmax) : kmax;
        uint64_t done = 0;
        while (done < (uint64_t)total) {
            const int64_t n = std::min<int64_t>(K, (int64_t)((uint64_t)total - done));
            ra.mcs0 = (uint32_t)(base + done);
            ra.nmcs = (int)n;
            ra.mail = h->d_res_mail;
            // the state a failed launch is redone from, cleared mailboxes,
            // error words and boxes (one kernel)
            gcpm::ResPrep rp{};
            rp.lat = cur_lat;
            rp.lat_backup = h->d_backup_lat;
            rp.lat_bytes = lat_bytes;
            rp.mail = h->d_res_mail;
            rp.nmail = nmail;
            rp.tab = h->d_tab[0];
            rp.tab_backup = h->d_backup_tab;
            rp.bbox = h->d_bbox;
            rp.rows = (int)rows;
            rp.flags = h->d_res_flags;
            rp.counters = h->counters_on ? h->d_counters : nullptr;
            HIP_TRY(gcpm::launch_res_prep(rp, h->stream));
            HIP_TRY(gcpm::launch_resident_guard(bits != 0, ra.k, h->d_anchor, h->d_bbox, ra.lte, ra.lnt,
                                                h->d_res_flags + 2, bad, h->stream));
            hipEvent_t e0, e1;
            if (int rc = h->next_events(&e0, &e1)) return rc;
            if (e0) HIP_TRY(hipEventRecord(e0, h->stream));
            HIP_TRY(gcpm::launch_resident(bits != 0, res_var, ra, tiles, h->stream));
            if (e1) HIP_TRY(hipEventRecord(e1, h->stream));
            unsigned int fl[2] = {0, 0};
            HIP_TRY(hipMemcpyAsync(fl, err, sizeof(fl), hipMemcpyDeviceToHost, h->stream));
            HIP_TRY(hipStreamSynchronize(h->stream));
            if (bits) h->wide_valid = false;
            if (fl[0] || fl[1]) {
                // The launch's result does not stand: restore its start and
                // make its MCS with one launch per phase.
                HIP_TRY(hipMemcpyAsync(cur_lat, h->d_backup_lat, lat_bytes, hipMemcpyDeviceToDevice, h->stream));
                for (int i = 0; i < 2; i++)
                    HIP_TRY(hipMemcpyAsync(h->d_tab[i], h->d_backup_tab, sizeof(int64_t) * rows,
                                           hipMemcpyDeviceToDevice, h->stream));
                if (h->counters_on)  // the MCS are counted when they are redone
                    HIP_TRY(hipMemcpyAsync(h->d_counters, h->d_counters + 4, 4 * sizeof(unsigned long long),
                                           hipMemcpyDeviceToDevice, h->stream));
                HIP_TRY(hipStreamSynchronize(h->stream));
                if (fl[1]) {
                    h->res_fallbacks++;
                    // Too many cells near a 64^2 tile for its slots, or
                    // shared with one neighbour, and none wider than two
                    // tiles: tiles of 32^2 meet a quarter as many cells and
                    // half as much border, so the rest of the run (and the
                    // lattice's later runs) take those, unless an edge was
                    // asked for or they do not fit co-resident.
                    int alt = 0;
                    if (res_lte == 6 && h->want_tile == 0 && !(fl[1] & gcpm::kResGuardWide) &&
                        (fl[1] & (gcpm::kResGuardSlots | gcpm::kResGuardShared)) != 0) {
                        int maxb = 0;
                        HIP_TRY(gcpm::resident_max_blocks(bits != 0, 5, res_var, h->n, &maxb));
                        if ((1 << (2 * (h->lgL - 5))) <= maxb) alt = 5;
                    }
                    if (alt) {
                        h->res_small_tiles = true;
                        res_lte = alt;
                        tiles = 1 << (2 * (h->lgL - res_lte));
                        plan.blocks = tiles;
                        plan.ltile = res_lte;
                        ra.lte = res_lte;
                        ra.lnt = h->lgL - res_lte;
                        if (int rc = mail_for()) return rc;
                    } else {
                        // a cell too wide, too many cells for a tile, or
                        // misses even in the shortest launches: the
                        // automatic choice stops taking the form for this
                        // lattice
                        if (h->want_form == -1 &&
                            ((fl[1] & ~gcpm::kResGuardMiss) != 0 || n <= kResMinMcs))
                            h->res_auto_off = true;
                        K = std::max<int64_t>(kResMinMcs, n / 2);
                        h->res_mcs = K;
                        h->res_streak = 0;
                    }
                } else {
                    h->res_wait_fallbacks++;
                }
                h->mcs = base + done;
                cpm_run_args part = *a;
                part.iterations = (int)n;
                part.inner_iterations = (int)n;
                const int saved = h->want_form;
                h->want_form = Plan::kPhase;
                const int rc = cpm_run(h, &part);
                h->want_form = saved;
                if (rc) return rc;
                // (the per-phase launches leave the 32-bit or the 16-bit copy
                // current, as any run; the next launch converts as needed)
                if (bits) {
                    if (int rc2 = h->ensure_narrow(bits)) return rc2;
                } else {
                    if (int rc2 = h->ensure_wide()) return rc2;
                }
            } else {
                h->last_plan = plan;
                if (h->timing) {
                    h->t_launches++;
                    h->t_updates += updates_per_mcs * n;
                }
                if (K < kmax && ++h->res_streak >= 4) {
                    K = std::min<int64_t>(2 * K, kmax);
                    h->res_mcs = K;
                    h->res_streak = 0;
                }
            }
            done += (uint64_t)n;
        }
        h->small_narrow = bits != 0;
        h->res_var = res_var;
        h->mcs = base + (uint64_t)total;
        return CPM_OK;
    }

    if (plan.mode == Plan::kSmall) {
        gcpm::SmallArgs sa{};
        sa.k = k;
        sa.lh = lh;
        sa.nmcs = a->inner_iterations;
        const int bits = h->narrow_bits();
        if (bits) {
            if (int rc = h->ensure_narrow(bits)) return rc;
            sa.k.lat16 = h->d_lat16;
            sa.k.idbits = (uint32_t)bits;
            sa.k.idmask = (1u << bits) - 1u;
        } else {
            if (int rc = h->ensure_wide()) return rc;
            h->narrow_valid = false;
        }
        // the thresholds go to LDS too when they fit beside the lattice
        sa.thr_lds = gcpm::small_lds_bytes(h->d, h->lgL, h->n, bits != 0) + 4 * (size_t)k.nthr <= kSmallLdsBudget;
        const bool pers = k.pers != 0;
        if (pers && !h->ring_valid)
            if (int rc = h->init_ring()) return rc;
        for (long ch = 0; ch < chunks; ch++) {
            sa.mcs0 = (uint32_t)(h->mcs + (uint64_t)ch * a->inner_iterations);
            hipEvent_t e0, e1;
            if (int rc = h->next_events(&e0, &e1)) return rc;
            HIP_TRY(gcpm::launch_small(h->d, bits != 0, sa, h->stream, e0, e1));
            if (bits) h->wide_valid = false;
            if (h->timing) {
                h->t_launches++;
                h->t_updates += updates_per_mcs * a->inner_iterations;
            }
            // end of a chunk (one reference launch): the tables are settled;
            // update the preferred directions from the centroid history
            if (pers)
                if (int rc = h->update_pref_dirs()) return rc;
        }
        h->small_narrow = bits != 0;
        h->mcs += (uint64_t)total;
        return finish_run_tune();
    }

    if (plan.mode == Plan::kPhase) {
        gcpm::PhaseArgs pa{};
        pa.k = k;
        pa.lh = lh;
        pa.nsub = 1 << (lh * h->d);
        // box3 stages 3D boxes in LDS (c = 1).  Measured slower than global
        // reads at 256^3 (61 vs 44 us per phase), so it is opt-in.
        pa.box3 = pick(h->opts.box3, 0);
        // The per-phase kernels run on the 16-bit lattice when every site
        // fits (the LDS box form reads 32-bit sites only).
        const int bits = pa.box3 ? 0 : h->narrow_bits();
        if (bits) {
            if (int rc = h->ensure_narrow(bits)) return rc;
            pa.k.lat16 = h->d_lat16;
            pa.k.idbits = (uint32_t)bits;
            pa.k.idmask = (1u << bits) - 1u;
        } else {
            if (int rc = h->ensure_wide()) return rc;
            h->narrow_valid = false;
        }
        // Kernel form (defaults measured fastest): small reads parameters
        // without LDS staging (default in 3D), compact = 0 finishes every
        // attempt in its own thread, full queues whole neighbourhoods, att
        // attempts are drawn per thread by the compacted kernel (8 in 3D, 2
        // in 2D, where the pair shares one Philox block).
        const KernelOpts& o = h->opts;
        pa.form = (pick(o.small, h->d == 3) ? gcpm::kFormSmall : 0) |
                  (pick(o.compact, 1) ? gcpm::kFormCompact : 0) |
                  (pick(o.full, 0) ? gcpm::kFormFull : 0) |
                  gcpm::att_form(pick(o.att, h->d == 3 ? 8 : 2));
        pa.lstride = k.lstride;
        // Autotuner: the candidate forms of this workload (all compute the
        // same results; the special launches -- cell_sync, split phases, the
        // reference stream, the LDS box form -- are not tuned).
        FormTuner::Entry* tune = nullptr;
        if (h->autotune && !a->cell_sync && !split && !k.refstream && !pa.box3 && !h->timing) {
            const bool sm = h->nt <= gcpm::kSmallTypes;
            std::vector<int> forms;
            const int C = gcpm::kFormCompact, SM = gcpm::kFormSmall;
            if (u == 1 && h->d == 2) {
                forms = {C, C | gcpm::kFormAtt2, C | gcpm::kFormAtt4};
                if (sm) forms.push_back(C | SM);
            } else if (u == 1) {
                forms = {C};
                if (sm) forms.insert(forms.end(), {C | SM | gcpm::kFormAtt4, C | SM | gcpm::kFormAtt8});
            } else {
                forms = {0};
                if (sm) forms.push_back(SM);
            }
            const uint64_t key = (uint64_t)1 << 63 | (uint64_t)h->d | (uint64_t)h->lgL << 2 | (uint64_t)ilog2(c) << 8 |
                                 (uint64_t)u << 12 | (uint64_t)k.roles << 24 | (uint64_t)sm << 25 |
                                 (uint64_t)(bits != 0) << 26 | (uint64_t)k.chem << 27 |
                                 (uint64_t)k.act << 28 | (uint64_t)k.pers << 29 | (uint64_t)k.S << 32;
            tune = &h->tuner.entry(key, forms);
            if (tune->best >= 0) {
                pa.form = tune->best;
                tune = nullptr;
            }
        }
        const int form0 = pa.form;
        int g = 0;  // phase counter since the tables were last settled
        const bool pers = k.pers != 0;
        if (pers && !h->ring_valid)
            if (int rc = h->init_ring()) return rc;
        for (long m = 0; m < total; m++) {
            pa.mcs = (uint32_t)(h->mcs + (uint64_t)m);
            if (bits) h->wide_valid = false;  // (a chunk end may have refreshed it)
            // The reference seeds its threads at every launch of
            // inner_iterations MCS, from the same seed (cpm.cu:347-360).
            if (k.refstream && m % a->inner_iterations == 0)
                HIP_TRY(gcpm::launch_ref_seed(h->d, k, (uint32_t)h->seed, h->stream));
            for (int r = 0; r < k.rounds; r++) {
                for (int pc = 0; pc < k.phases_per_round; pc++, g++) {
                    phase_offsets(h->d, k, pc, &pa.cx, &pa.cy, &pa.cz);
                    phase_slice(h->d, k, pc, &pa.gsx, &pa.gsy, &pa.gsz);
                    pa.phase = (uint32_t)(r * k.phases_per_round + pc);
                    pa.g = g;
                    // cell_sync: per attempt index, a claim launch and a commit
                    // launch; otherwise one launch makes all u attempts
                    const int nidx = a->cell_sync ? u : 1;
                    for (int i = 0; i < nidx; i++) {
                        pa.att0 = a->cell_sync ? i : 0;
                        pa.natt = a->cell_sync ? 1 : u;
                        if (a->cell_sync) {
                            // the tag only grows: claims never need clearing
                            if (++h->sync_tag == 0) return fail(CPM_EINVAL, "cell_sync phase tag overflow");
                            pa.sync_tag = h->sync_tag;
                            pa.claims = h->d_claims;
                        }
                        // split: decide (with cell_sync also claim), then
                        // apply; cell_sync alone: claim, then commit
                        if (split || a->cell_sync) {
                            pa.decide = h->d_decide;
                            pa.sync_pass = split ? (a->cell_sync ? 5 : 3) : 1;
                            HIP_TRY(gcpm::launch_phase(h->d, pa, h->stream, nullptr, nullptr));
                            pa.sync_pass = split ? (a->cell_sync ? 6 : 4) : 2;
                        }
                        hipEvent_t e0, e1;
                        if (int rc = h->next_events(&e0, &e1)) return rc;
                        const int ti = tune ? h->tuner.pick(*tune) : -1;
                        if (ti >= 0) {
                            pa.form = tune->forms[ti];
                            HIP_TRY(h->tuner.events(&e0, &e1));
                        }
                        HIP_TRY(gcpm::launch_phase(h->d, pa, h->stream, e0, e1, &h->last_kernel));
                        if (ti >= 0) h->tuner.pending.push_back({tune, ti, e0, e1});
                        pa.form = form0;
                        if (h->timing) {
                            h->t_launches++;
                            h->t_updates += (int64_t)pa.nsub * pa.natt;
                        }
                    }
                }
            }
            g &= 1;  // only the parity matters
            if (pers && (m + 1) % a->inner_iterations == 0) {
                // end of a chunk (one reference launch): settle, then update
                // the preferred directions from the centroid history
                HIP_TRY(gcpm::launch_settle(h->d_tab, h->d_del, g, h->n, h->stream));
                g = 0;
                if (int rc = h->update_pref_dirs()) return rc;
            }
        }
        // Fold the last phase's deltas into the lagging base: both bases
        // then hold the current state and both delta buffers are zero.
        HIP_TRY(gcpm::launch_settle(h->d_tab, h->d_del, g, h->n, h->stream));
        if (int rc = finish_run_tune()) return rc;
        HIP_TRY(h->tuner.settle());
    } else {
        if (int rc = h->ensure_wide()) return rc;
        h->narrow_valid = false;
        gcpm::PersistArgs pa{};
        pa.k = k;
        pa.lh = lh;
        pa.bar = h->d_bar;
        pa.ltx = plan.ltile;
        pa.lty = plan.ltile;
        pa.lntx = h->lgL - plan.ltile;
        pa.nmcs = a->inner_iterations;
        const bool lds = plan.mode == Plan::kLds2d;
        for (long ch = 0; ch < chunks; ch++) {
            pa.mcs0 = (uint32_t)(h->mcs + (uint64_t)ch * a->inner_iterations);
            hipEvent_t e0, e1;
            if (int rc = h->next_events(&e0, &e1)) return rc;
            if (e0) HIP_TRY(hipEventRecord(e0, h->stream));
            HIP_TRY(gcpm::launch_persist(h->d, lds, pa, plan.blocks, plan.lds_bytes, h->stream));
            if (e1) HIP_TRY(hipEventRecord(e1, h->stream));
            if (h->timing) {
                h->t_launches++;
                h->t_updates += updates_per_mcs * a->inner_iterations;
            }
        }
        if (int rc = check_barrier(h->d_bar, 3, h->stream)) return rc;
    }
    h->mcs += (uint64_t)total;
    return CPM_OK;
}

int cpm_sync(cpm_handle h) {
    if (int rc = check_handle(h)) return rc;
    HIP_TRY(hipStreamSynchronize(h->stream));
    return CPM_OK;
}

int cpm_pull(cpm_handle h) {
    if (int rc = check_handle(h)) return rc;
    if (int rc = h->ensure_wide()) return rc;
    HIP_TRY(hipMemcpyAsync(h->h_lat, h->d_lat, h->sites * 4, hipMemcpyDeviceToHost, h->stream));
    if (h->has_act)  // cpm.cu:535
        HIP_TRY(hipMemcpyAsync(h->h_act, h->d_act, h->sites * 4, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    // cpm.cu:537-538 copies the device field back: the sweep never writes
    // it, so that is the field as last pushed.
    if (h->has_field) h->h_field = h->h_field_pushed;
    return CPM_OK;
}

int cpm_get_cell_table(cpm_handle h, int32_t* area, int32_t* perim, int32_t* anchor, int64_t* csum) {
    if (int rc = check_handle(h)) return rc;
    if (!h->pushed) return fail(CPM_ESTATE, "cell tables exist only after push_to_gpu");
    const size_t rows = (size_t)h->n + 1;
    std::vector<int64_t> tab(rows);
    HIP_TRY(hipMemcpyAsync(tab.data(), h->d_tab[0], rows * 8, hipMemcpyDeviceToHost, h->stream));
    if (anchor) HIP_TRY(hipMemcpyAsync(anchor, h->d_anchor, rows * 3 * 4, hipMemcpyDeviceToHost, h->stream));
    if (csum) {
        // Centroid sums about the push-time anchors, rebuilt from the current
        // lattice: equal to sums kept per accepted flip (each site adds
        // wrap(x - anchor) while it belongs to the cell).
        if (int rc = h->ensure_wide()) return rc;
        HIP_TRY(hipMemsetAsync(h->d_csum, 0, rows * 3 * 8, h->stream));
        HIP_TRY(gcpm::launch_csum(h->d, h->d_lat, h->lgL, h->d_anchor, h->d_csum, h->stream));
        HIP_TRY(hipMemcpyAsync(csum, h->d_csum, rows * 3 * 8, hipMemcpyDeviceToHost, h->stream));
    }
    HIP_TRY(hipStreamSynchronize(h->stream));
    for (size_t c = 0; c < rows; c++) {
        if (area) area[c] = (int32_t)(uint32_t)tab[c];
        if (perim) perim[c] = (int32_t)(tab[c] >> 32);
    }
    return CPM_OK;
}

int cpm_get_centroids(cpm_handle h, double* out, int capacity, int* rows) {
    if (int rc = check_handle(h)) return rc;
    if (rows) *rows = h->active;
    if (!h->pushed) return fail(CPM_ESTATE, "centroids exist only after push_to_gpu");
    if (capacity < h->active || (!out && h->active)) return fail(CPM_EINVAL, "output too small");
    const size_t nr = (size_t)h->n + 1;
    std::vector<int32_t> area(nr), anchor(nr * 3);
    std::vector<int64_t> csum(nr * 3);
    if (int rc = cpm_get_cell_table(h, area.data(), nullptr, anchor.data(), csum.data())) return rc;
    // cpm.cu:68-110 reports sum/count per axis for cells 1..activeCells.
    for (int c = 1; c <= h->active; c++) {
        for (int ax = 0; ax < h->d; ax++) {
            double v;
            if (!area[c]) {
                v = NAN;
            } else {
                v = (double)anchor[3 * c + ax] + (double)csum[3 * c + ax] / (double)area[c];
                if (v < 0.0) v += h->L;
                if (v >= h->L) v -= h->L;
            }
            out[(size_t)(c - 1) * h->d + ax] = v;
        }
    }
    return CPM_OK;
}

int cpm_timing_begin(cpm_handle h) {
    if (int rc = check_handle(h)) return rc;
    h->timing = true;
    h->ev_used = 0;
    h->t_launches = 0;
    h->t_updates = 0;
    return CPM_OK;
}

int cpm_timing_end(cpm_handle h, double* kernel_ms, int64_t* launches, int64_t* site_updates) {
    if (int rc = check_handle(h)) return rc;
    HIP_TRY(hipStreamSynchronize(h->stream));
    double total = 0.0;
    for (size_t i = 0; i + 1 < h->ev_used; i += 2) {
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, h->ev[i], h->ev[i + 1]));
        total += ms;
    }
    if (kernel_ms) *kernel_ms = total;
    if (launches) *launches = h->t_launches;
    if (site_updates) *site_updates = h->t_updates;
    h->timing = false;
    h->ev_used = 0;
    return CPM_OK;
}

int cpm_last_launch(cpm_handle h, int* form, int* blocks, int* tile) {
    if (int rc = check_handle(h)) return rc;
    if (form) *form = (int)h->last_plan.mode;
    if (blocks) *blocks = h->last_plan.blocks;
    if (tile) *tile = h->last_plan.ltile ? 1 << h->last_plan.ltile : 0;
    return CPM_OK;
}

const char* cpm_last_kernel(cpm_handle h) {
    if (check_handle(h)) return "";
    switch (h->last_plan.mode) {
        case Plan::kGlobal: return h->d == 3 ? "k_persist_global<3>" : "k_persist_global<2>";
        case Plan::kLds2d: return "k_persist_lds2d";
        case Plan::kResident: {
            static const char* const names[2][2][4] = {
                {{"k_resident2d<unsigned int, 5, 0>", "k_resident2d<unsigned int, 5, 1>",
                  "k_resident2d<unsigned int, 5, 2>", "k_resident2d<unsigned int, 5, 3>"},
                 {"k_resident2d<unsigned int, 6, 0>", "k_resident2d<unsigned int, 6, 1>",
                  "k_resident2d<unsigned int, 6, 2>", "k_resident2d<unsigned int, 6, 3>"}},
                {{"k_resident2d<unsigned short, 5, 0>", "k_resident2d<unsigned short, 5, 1>",
                  "k_resident2d<unsigned short, 5, 2>", "k_resident2d<unsigned short, 5, 3>"},
                 {"k_resident2d<unsigned short, 6, 0>", "k_resident2d<unsigned short, 6, 1>",
                  "k_resident2d<unsigned short, 6, 2>", "k_resident2d<unsigned short, 6, 3>"}}};
            return names[h->small_narrow][h->last_plan.ltile == 6][h->res_var];
        }
        case Plan::kSmall:
            return h->d == 3 ? (h->small_narrow ? "k_small_lds<3, unsigned short>" : "k_small_lds<3, unsigned int>")
                             : (h->small_narrow ? "k_small_lds<2, unsigned short>" : "k_small_lds<2, unsigned int>");
        default: return gcpm::phase_kernel_name(h->last_kernel);
    }
}

int cpm_set_launch(cpm_handle h, int form, int tile) {
    if (int rc = check_handle(h)) return rc;
    if (form < -1 || form > 4) return fail(CPM_EINVAL, "form must be -1 (auto), 0, 1, 2, 3 or 4");
    if (tile < 0 || (tile && !is_pow2(tile))) return fail(CPM_EINVAL, "tile must be 0 or a power of two");
    h->want_form = form;
    h->want_tile = tile;
    return CPM_OK;
}

int cpm_set_kernel_options(cpm_handle h, int small, int compact, int full, int attempts,
                           int narrow, int box3) {
    if (int rc = check_handle(h)) return rc;
    for (int v : {small, compact, full, narrow, box3})
        if (v < -1 || v > 1) return fail(CPM_EINVAL, "kernel options must be -1 (default), 0 or 1");
    if (attempts != -1 && attempts != 1 && attempts != 2 && attempts != 4 && attempts != 8)
        return fail(CPM_EINVAL, "attempts per thread must be -1 (default), 1, 2, 4 or 8");
    h->opts = KernelOpts{small, compact, full, attempts, narrow, box3};
    return CPM_OK;
}

int cpm_resident_fallbacks(cpm_handle h, int64_t* guard, int64_t* waits) {
    if (int rc = check_handle(h)) return rc;
    if (guard) *guard = h->res_fallbacks;
    if (waits) *waits = h->res_wait_fallbacks;
    return CPM_OK;
}

int cpm_set_counters(cpm_handle h, int on) {
    if (int rc = check_handle(h)) return rc;
    if (on && !h->d_counters) HIP_TRY(hipMalloc(&h->d_counters, 8 * sizeof(unsigned long long)));
    if (on) {
        HIP_TRY(hipMemsetAsync(h->d_counters, 0, 4 * sizeof(unsigned long long), h->stream));
        h->ctr_mcs_acc = 0;
        h->ctr_mcs0 = h->mcs;
    } else if (h->counters_on) {
        h->ctr_mcs_acc += h->mcs - h->ctr_mcs0;
    }
    h->counters_on = on != 0;
    return CPM_OK;
}

int cpm_get_counters(cpm_handle h, int64_t out[5], int reset) {
    if (int rc = check_handle(h)) return rc;
    if (!out) return fail(CPM_EINVAL, "null output");
    if (!h->d_counters) return fail(CPM_ESTATE, "counters were never switched on (cpm_set_counters)");
    unsigned long long c[4];
    HIP_TRY(hipMemcpyAsync(c, h->d_counters, sizeof(c), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    const uint64_t mcs = h->ctr_mcs_acc + (h->counters_on ? h->mcs - h->ctr_mcs0 : 0);
    out[0] = (int64_t)(mcs * (uint64_t)h->sites);
    for (int i = 0; i < 4; i++) out[1 + i] = (int64_t)c[i];
    if (reset) {
        HIP_TRY(hipMemsetAsync(h->d_counters, 0, sizeof(c), h->stream));
        h->ctr_mcs_acc = 0;
        h->ctr_mcs0 = h->mcs;
    }
    return CPM_OK;
}

int cpm_set_autotune(cpm_handle h, int on) {
    if (int rc = check_handle(h)) return rc;
    h->autotune = on != 0;
    return CPM_OK;
}

int cpm_set_rng(cpm_handle h, int mode) {
    if (int rc = check_handle(h)) return rc;
    if (mode != 0 && mode != 1) return fail(CPM_EINVAL, "rng mode must be 0 (counter stream) or 1 (reference stream)");
    if (mode != h->rng) h->params_dirty = true;  // the acceptance table differs
    h->rng = mode;
    return CPM_OK;
}

int cpm_get_rng(cpm_handle h, int* mode) {
    if (int rc = check_handle(h)) return rc;
    if (!mode) return fail(CPM_EINVAL, "null output");
    *mode = h->rng;
    return CPM_OK;
}

int cpm_set_seed(cpm_handle h, uint64_t seed) {
    if (int rc = check_handle(h)) return rc;
    h->seed = seed;
    h->init_pref_dirs();  // the initial preferred directions are drawn from the seed
    return CPM_OK;
}

int cpm_get_seed(cpm_handle h, uint64_t* seed) {
    if (int rc = check_handle(h)) return rc;
    if (!seed) return fail(CPM_EINVAL, "null output");
    *seed = h->seed;
    return CPM_OK;
}

int cpm_set_mcs(cpm_handle h, uint64_t mcs) {
    if (int rc = check_handle(h)) return rc;
    h->mcs = mcs;
    return CPM_OK;
}

int cpm_get_mcs(cpm_handle h, uint64_t* mcs) {
    if (int rc = check_handle(h)) return rc;
    if (!mcs) return fail(CPM_EINVAL, "null output");
    *mcs = h->mcs;
    return CPM_OK;
}

}  // extern "C"

namespace {

// Device buffers and stream of the batched statistics launch
// (cpm_energy_stats_many), shared by every engine of the process and kept
// until it ends; calls are serialized.
struct StatsRunner {
    std::mutex mu;
    hipStream_t stream = nullptr;
    hipEvent_t ready = nullptr, e0 = nullptr, e1 = nullptr;
    unsigned long long* d_out = nullptr;  // [engines][500]
    unsigned long long* h_out = nullptr;  // pinned copy
    gcpm::StatsDesc* d_desc = nullptr;
    int* d_map = nullptr;
    size_t cap = 0, map_cap = 0;
    double last_ms = 0.0;  // kernel time of the last call

    int ensure(int n) {
        if (!stream) {
            HIP_TRY(hipStreamCreate(&stream));
            HIP_TRY(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
            HIP_TRY(hipEventCreate(&e0));
            HIP_TRY(hipEventCreate(&e1));
        }
        if ((size_t)n > cap) {
            constexpr size_t nw = CPM_NTYPES * CPM_NTYPES + CPM_NTYPES * 5;
            (void)hipFree(d_out);
            (void)hipFree(d_desc);
            if (h_out) (void)hipHostFree(h_out);
            d_out = nullptr;
            d_desc = nullptr;
            h_out = nullptr;
            cap = 0;
            HIP_TRY(hipMalloc(&d_out, sizeof(unsigned long long) * nw * n));
            HIP_TRY(hipMalloc(&d_desc, sizeof(gcpm::StatsDesc) * n));
            HIP_TRY(hipHostMalloc((void**)&h_out, sizeof(unsigned long long) * nw * n, hipHostMallocDefault));
            cap = (size_t)n;
        }
        return CPM_OK;
    }
    int ensure_map(size_t blocks) {
        if (blocks > map_cap) {
            (void)hipFree(d_map);
            d_map = nullptr;
            map_cap = 0;
            HIP_TRY(hipMalloc(&d_map, sizeof(int) * blocks));
            map_cap = blocks;
        }
        return CPM_OK;
    }
};

StatsRunner& stats_runner() {
    static StatsRunner* r = new StatsRunner();  // never destroyed: it may outlive the HIP runtime
    return *r;
}

// Batched run, one launch per colour phase for all replicas (every replica
// has the same c and u, hence the same phase sequence).  split: every phase
// is a decide launch and an apply launch (as cpm_run with the act term,
// shared roles and c = 1), on the uncompacted layout.
int batch_run_phase(cpm_batch* b, const cpm_run_args* a, std::vector<gcpm::ReplicaDesc>& reps,
                    long total, bool split, bool graph) {
    const int n = (int)b->sims.size();
    uint32_t tag = 0;  // cell_sync: one tag for all replicas, above each one's earlier tags
    bool pers = false;
    for (int i = 0; i < n; i++) {
        cpm_handle h = b->sims[i];
        if (a->cell_sync && !h->d_claims) {
            HIP_TRY(hipMalloc(&h->d_claims, sizeof(uint64_t) * ((size_t)h->n + 1)));
            HIP_TRY(hipMemsetAsync(h->d_claims, 0, sizeof(uint64_t) * ((size_t)h->n + 1), h->stream));
        }
        reps[i].claims = h->d_claims;
        tag = std::max(tag, h->sync_tag);
        if (reps[i].k.pers) {
            pers = true;
            if (!h->ring_valid)
                if (int rc = h->init_ring()) return rc;
        }
    }
    // Kernel form, as for single runs (cpm_batch_set_kernel_options):
    // compacted when u = 1, parameters without staging when every replica
    // has at most four types (default in 3D), att attempts drawn per thread.
    const int u1 = a->updates_per_checkerboard_switch;
    const bool refstream = reps[0].k.refstream != 0;  // (all replicas alike)
    const bool compact = u1 == 1 && !refstream && !split && pick(b->opts.compact, 1) != 0;
    bool small = pick(b->opts.small, b->d == 3) != 0;
    for (int i = 0; i < n; i++) small = small && reps[i].k.nt <= gcpm::kSmallTypes;
    // Attempts drawn per thread: 1, 4 or 8 (a batch is throughput bound, so
    // fewer, fuller blocks stage the parameters fewer times).
    int att = compact ? pick(b->opts.att, b->d == 3 ? 8 : 4) : 1;
    att = att >= 8 ? 8 : att >= 4 ? 4 : 1;
    const int form_default = (compact ? gcpm::kFormCompact : 0) | (small ? gcpm::kFormSmall : 0) |
                             gcpm::att_form(att);
    // 16-bit lattices when every replica's sites fit, else all 32-bit.
    bool narrow = true;
    for (int i = 0; i < n; i++) narrow = narrow && b->sims[i]->narrow_bits() != 0;
    for (int i = 0; i < n; i++) {
        cpm_handle h = b->sims[i];
        if (narrow) {
            const int bits = h->narrow_bits();
            if (int rc = h->ensure_narrow(bits)) return rc;
            reps[i].k.lat16 = h->d_lat16;
            reps[i].k.idbits = (uint32_t)bits;
            reps[i].k.idmask = (1u << bits) - 1u;
        } else {
            if (int rc = h->ensure_wide()) return rc;
            h->narrow_valid = false;
        }
    }
    // Autotuner (cpm_batch_set_autotune): attempts per thread 1 / 4 / 8, with
    // staged or (at most four types) scalar parameters; every form computes
    // the same results.
    FormTuner::Entry* tune = nullptr;
    int form = form_default;
    if (b->autotune && compact && !a->cell_sync) {
        bool sm = true;
        for (int i = 0; i < n; i++) sm = sm && reps[i].k.nt <= gcpm::kSmallTypes;
        std::vector<int> forms;
        for (int t : {1, 4, 8}) {
            forms.push_back(gcpm::kFormCompact | gcpm::att_form(t));
            if (sm) forms.push_back(gcpm::kFormCompact | gcpm::kFormSmall | gcpm::att_form(t));
        }
        uint64_t key = (uint64_t)b->d | (uint64_t)n << 2 | (uint64_t)reps[0].k.lc << 20 |
                       (uint64_t)reps[0].k.roles << 24 | (uint64_t)narrow << 25 | (uint64_t)sm << 26;
        for (int i = 0; i < n; i++) key = key * 1000003u + (uint64_t)reps[i].lh;
        tune = &b->tuner.entry(key, forms);
        if (tune->best >= 0) {
            form = tune->best;
            tune = nullptr;
        }
    }
    // Block layouts: each replica's blocks cover its sub-blocks, att x 256
    // per block when compacted (one layout per attempts-per-thread value in
    // use, stacked in d_reps / d_rep_of).
    std::vector<int> atts;
    if (tune) atts = {1, 4, 8};
    else atts = {compact ? gcpm::form_att(form) : 1};
    std::vector<gcpm::ReplicaDesc> all_reps;
    std::vector<int> rep_of, lay_blocks, lay_rep_of0;
    for (int t : atts) {
        const long per_block = (long)gcpm::kBlock * t;
        int nb_total = 0;
        lay_rep_of0.push_back((int)rep_of.size());
        for (int i = 0; i < n; i++) {
            const long nsub = 1L << (reps[i].lh * b->d);
            // compacted: the fold strides over the cells, so blocks follow the draws
            const long nb = compact ? std::max(1L, (nsub + per_block - 1) / per_block)
                                    : (std::max(nsub, (long)b->sims[i]->n + 1) + gcpm::kBlock - 1) / gcpm::kBlock;
            reps[i].block0 = nb_total;
            reps[i].nblocks = (int)nb;
            reps[i].mcs0 = (uint32_t)b->sims[i]->mcs;
            nb_total += reps[i].nblocks;
            for (int j = 0; j < reps[i].nblocks; j++) rep_of.push_back(i);
        }
        lay_blocks.push_back(nb_total);
        all_reps.insert(all_reps.end(), reps.begin(), reps.end());
    }
    const int total_blocks = (int)rep_of.size();
    if (total_blocks > b->rep_of_cap) {
        (void)hipFree(b->d_rep_of);
        b->d_rep_of = nullptr;
        HIP_TRY(hipMalloc(&b->d_rep_of, sizeof(int) * total_blocks));
        b->rep_of_cap = total_blocks;
    }
    auto layout_of = [&](int f) {
        const int t = compact ? gcpm::form_att(f) : 1;
        for (size_t li = 0; li < atts.size(); li++)
            if (atts[li] == t) return (int)li;
        return 0;
    };
    const int lay0 = layout_of(form);
    int nblocks = lay_blocks[lay0];
    b->last_blocks = nblocks;
    for (int i = 0; i < n; i++) {
        HIP_TRY(hipEventRecord(b->ready[i], b->sims[i]->stream));
        HIP_TRY(hipStreamWaitEvent(b->stream, b->ready[i], 0));
    }
    HIP_TRY(hipMemcpyAsync(b->d_rep_of, rep_of.data(), sizeof(int) * total_blocks, hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipMemcpyAsync(b->d_reps, all_reps.data(), sizeof(gcpm::ReplicaDesc) * all_reps.size(),
                           hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));  // the layouts are host temporaries
    gcpm::BatchPhaseArgs pa{};
    auto set_layout = [&](int li) {
        pa.reps = b->d_reps + (size_t)li * n;
        pa.rep_of = b->d_rep_of + lay_rep_of0[li];
        nblocks = lay_blocks[li];
    };
    set_layout(lay0);
    pa.form = form;
    pa.narrow = narrow ? 1 : 0;
    pa.refstream = refstream ? 1 : 0;
    const gcpm::SweepConsts& k0 = reps[0].k;
    const int rounds = k0.rounds;
    pa.lstride = k0.lstride;
    const int u = a->updates_per_checkerboard_switch;
    int64_t attempts_per_launch = 0;  // sub-blocks of one colour phase over all replicas
    for (int i = 0; i < n; i++) attempts_per_launch += (int64_t)1 << (reps[i].lh * b->d);
    int g = 0;
    // Graph replay (batch form 2): the launches of G MCS (G the largest
    // divisor of the run's MCS up to kGraphMcs) are captured once as a hip
    // graph, with the MCS offset read from a device counter that a last node
    // advances by G, and the graph is replayed total / G times (and kept for
    // later runs of the same shape).  Kernel nodes only: no memset nodes.
    // Not with cell_sync (its tags grow per launch), the reference stream
    // (reseeds per chunk), persistence (host updates per chunk), timing or
    // per-launch tuning.
    // The graph's phases keep the table parity g (which starts at 0), so the
    // phases of G MCS must be even in number.
    const long P = (long)rounds * k0.phases_per_round;
    long G = 0;
    for (long gg = std::min<long>(total, kGraphMcs); gg >= 1 && graph; gg--)
        if (total % gg == 0 && (gg * P) % 2 == 0) {
            G = gg;
            break;
        }
    const bool use_graph = G > 0 && !tune && !b->timing && !a->cell_sync && !refstream && !pers;
    if (use_graph) {
        uint64_t key = (uint64_t)G | (uint64_t)lay0 << 8 | (uint64_t)form << 12 | (uint64_t)narrow << 20 |
                       (uint64_t)split << 21 | (uint64_t)u << 22 | (uint64_t)rounds << 32 |
                       (uint64_t)k0.phases_per_round << 44 | (uint64_t)k0.lstride << 56;
        key = key * 1000003u + (uint64_t)nblocks;
        key = key * 1000003u + (uint64_t)(uintptr_t)pa.reps;
        key = key * 1000003u + (uint64_t)(uintptr_t)pa.rep_of;
        if (!b->gexec || b->gkey != key) {
            if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
            b->gexec = nullptr;
            HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            hipError_t ce = hipSuccess;
            pa.mbase = b->d_mbase;
            int gl = 0;
            for (long m = 0; m < G && ce == hipSuccess; m++) {
                pa.m = (uint32_t)m;
                for (int r = 0; r < rounds && ce == hipSuccess; r++) {
                    for (int pc = 0; pc < k0.phases_per_round && ce == hipSuccess; pc++, gl++) {
                        phase_offsets(b->d, k0, pc, &pa.cx, &pa.cy, &pa.cz);
                        phase_slice(b->d, k0, pc, &pa.gsx, &pa.gsy, &pa.gsz);
                        pa.phase = (uint32_t)(r * k0.phases_per_round + pc);
                        pa.g = gl & 1;
                        pa.att0 = 0;
                        pa.natt = u;
                        pa.sync_pass = 0;
                        if (split) {
                            pa.sync_pass = 3;
                            ce = gcpm::launch_batch_phase(b->d, pa, nblocks, b->stream);
                            pa.sync_pass = 4;
                        }
                        if (ce == hipSuccess)
                            ce = gcpm::launch_batch_phase(b->d, pa, nblocks, b->stream, nullptr, nullptr,
                                                          b->last_kernel, sizeof(b->last_kernel));
                    }
                }
            }
            if (ce == hipSuccess) ce = gcpm::launch_add_u32(b->d_mbase, (uint32_t)G, b->stream);
            hipGraph_t gr = nullptr;
            const hipError_t ee = hipStreamEndCapture(b->stream, &gr);
            pa.mbase = nullptr;
            if (ce != hipSuccess) {
                if (gr) (void)hipGraphDestroy(gr);
                return hip_fail(ce, "capturing the batch graph");
            }
            HIP_TRY(ee);
            const hipError_t ie = hipGraphInstantiate(&b->gexec, gr, nullptr, nullptr, 0);
            (void)hipGraphDestroy(gr);
            if (ie != hipSuccess) {
                b->gexec = nullptr;
                return hip_fail(ie, "hipGraphInstantiate (batch graph)");
            }
            b->gkey = key;
            b->graph_captures++;
        }
        HIP_TRY(gcpm::launch_set_u32(b->d_mbase, 0u, b->stream));
        for (long rep = 0; rep < total / G; rep++) HIP_TRY(hipGraphLaunch(b->gexec, b->stream));
        b->graph_replays += total / G;
        if (narrow)
            for (int i = 0; i < n; i++) b->sims[i]->wide_valid = false;
        // (an even number of phases per graph: g ends at 0)
    }
    b->last_form = use_graph ? 2 : 0;
    for (long m = 0; m < (use_graph ? 0 : total); m++) {
        pa.m = (uint32_t)m;
        if (narrow)  // (a chunk end may have refreshed them)
            for (int i = 0; i < n; i++) b->sims[i]->wide_valid = false;
        if (refstream && m % a->inner_iterations == 0)  // each chunk reseeds, as in cpm_run
            for (int i = 0; i < n; i++)
                HIP_TRY(gcpm::launch_ref_seed(b->d, reps[i].k, (uint32_t)b->sims[i]->seed, b->stream));
        for (int r = 0; r < rounds; r++) {
            for (int pc = 0; pc < k0.phases_per_round; pc++, g++) {
                phase_offsets(b->d, k0, pc, &pa.cx, &pa.cy, &pa.cz);
                phase_slice(b->d, k0, pc, &pa.gsx, &pa.gsy, &pa.gsz);
                pa.phase = (uint32_t)(r * k0.phases_per_round + pc);
                pa.g = g;
                // as cpm_run: cell_sync makes a claim and a commit launch per
                // attempt index
                const int nidx = a->cell_sync ? u : 1;
                for (int i = 0; i < nidx; i++) {
                    pa.att0 = a->cell_sync ? i : 0;
                    pa.natt = a->cell_sync ? 1 : u;
                    pa.sync_pass = 0;
                    if (a->cell_sync) {
                        if (++tag == 0) return fail(CPM_EINVAL, "cell_sync phase tag overflow");
                        pa.sync_tag = tag;
                    }
                    if (split || a->cell_sync) {
                        pa.sync_pass = split ? (a->cell_sync ? 5 : 3) : 1;
                        HIP_TRY(gcpm::launch_batch_phase(b->d, pa, nblocks, b->stream));
                        pa.sync_pass = split ? (a->cell_sync ? 6 : 4) : 2;
                    }
                    const int ti = tune ? b->tuner.pick(*tune) : -1;
                    hipEvent_t e0 = nullptr, e1 = nullptr;
                    if (ti >= 0) {
                        pa.form = tune->forms[ti];
                        set_layout(layout_of(pa.form));
                        HIP_TRY(b->tuner.events(&e0, &e1));
                    } else if (b->timing) {
                        if (int rc = b->next_events(&e0, &e1)) return rc;
                        b->t_launches++;
                        b->t_updates += attempts_per_launch * pa.natt;
                    }
                    HIP_TRY(gcpm::launch_batch_phase(b->d, pa, nblocks, b->stream, e0, e1, b->last_kernel,
                                                     sizeof(b->last_kernel)));
                    if (ti >= 0) {
                        b->tuner.pending.push_back({tune, ti, e0, e1});
                        pa.form = form;
                        set_layout(lay0);
                    }
                }
            }
        }
        g &= 1;
        if (pers && (m + 1) % a->inner_iterations == 0) {
            // end of a chunk: settle every replica, then update the preferred
            // directions of those with a persistence term (host, per replica)
            for (int i = 0; i < n; i++)
                HIP_TRY(gcpm::launch_settle(b->sims[i]->d_tab, b->sims[i]->d_del, g, b->sims[i]->n, b->stream));
            g = 0;
            HIP_TRY(hipStreamSynchronize(b->stream));
            for (int i = 0; i < n; i++)
                if (reps[i].k.pers)
                    if (int rc = b->sims[i]->update_pref_dirs()) return rc;
        }
    }
    for (int i = 0; i < n; i++) b->sims[i]->sync_tag = tag;
    HIP_TRY(b->tuner.settle());
    for (int i = 0; i < n; i++)
        HIP_TRY(gcpm::launch_settle(b->sims[i]->d_tab, b->sims[i]->d_del, g, b->sims[i]->n, b->stream));
    HIP_TRY(hipEventRecord(b->done, b->stream));
    for (int i = 0; i < n; i++) {
        HIP_TRY(hipStreamWaitEvent(b->sims[i]->stream, b->done, 0));
        b->sims[i]->mcs += (uint64_t)total;
    }
    return CPM_OK;
}

}  // namespace

extern "C" {

int cpm_batch_create(const cpm_handle* sims, int n, cpm_batch_handle* out) {
    if (!out) return fail(CPM_EINVAL, "null output handle");
    *out = nullptr;
    if (!sims || n < 1) return fail(CPM_EINVAL, "a batch needs at least one engine");
    for (int i = 0; i < n; i++) {
        if (!sims[i]) return fail(CPM_EINVAL, "null engine in batch");
        if (sims[i]->d != sims[0]->d) return fail(CPM_EINVAL, "all engines of a batch must have the same dimensionality");
        for (int j = 0; j < i; j++)
            if (sims[j] == sims[i]) return fail(CPM_EINVAL, "an engine appears twice in the batch");
    }
    cpm_batch* b = new (std::nothrow) cpm_batch();
    if (!b) return fail(CPM_EHIP, "host allocation failed");
    b->sims.assign(sims, sims + n);
    b->d = sims[0]->d;

    auto bail = [&](hipError_t err, const char* what) {
        int rc = hip_fail(err, what);
        delete b;
        return rc;
    };
    hipError_t err;
    if ((err = hipStreamCreate(&b->stream)) != hipSuccess) return bail(err, "hipStreamCreate");
    if ((err = hipEventCreateWithFlags(&b->done, hipEventDisableTiming)) != hipSuccess) return bail(err, "hipEventCreate");
    b->ready.resize(n, nullptr);
    for (int i = 0; i < n; i++)
        if ((err = hipEventCreateWithFlags(&b->ready[i], hipEventDisableTiming)) != hipSuccess)
            return bail(err, "hipEventCreate");
    if ((err = hipMalloc(&b->d_bars, sizeof(unsigned int) * (2 * n + 1))) != hipSuccess) return bail(err, "hipMalloc(barriers)");
    if ((err = hipMemset(b->d_bars, 0, sizeof(unsigned int) * (2 * n + 1))) != hipSuccess) return bail(err, "hipMemset(barriers)");
    // three block layouts at most (the autotuner's attempts-per-thread values)
    if ((err = hipMalloc(&b->d_reps, sizeof(gcpm::ReplicaDesc) * n * 3)) != hipSuccess)
        return bail(err, "hipMalloc(replicas)");
    if ((err = hipMalloc(&b->d_mbase, sizeof(uint32_t))) != hipSuccess) return bail(err, "hipMalloc(mcs offset)");
    *out = b;
    return CPM_OK;
}

int cpm_batch_destroy(cpm_batch_handle b) {
    delete b;
    return CPM_OK;
}

int cpm_batch_run(cpm_batch_handle b, const cpm_run_args* a) {
    if (!b) return fail(CPM_EINVAL, "null batch");
    const int n = (int)b->sims.size();
    std::vector<gcpm::ReplicaDesc> reps(n);
    std::vector<long> need(n);
    bool split = false;
    for (int i = 0; i < n; i++) {
        cpm_handle h = b->sims[i];
        if (h->rng != b->sims[0]->rng)
            return fail(CPM_EINVAL, "the engines of a batch must use the same random stream (cpm_set_rng)");
        if (int rc = prepare_run(h, a, &reps[i].k, &reps[i].lh)) return rc;
        if (h->rng == 1) {  // reference streams: one state per reference thread, as in cpm_run
            const size_t nreg = (size_t)1 << (b->d * reps[i].k.lr);
            if (nreg > h->rstate_cap) {
                (void)hipFree(h->d_rstate);
                h->d_rstate = nullptr;
                HIP_TRY(hipMalloc(&h->d_rstate, nreg * sizeof(int32_t)));
                h->rstate_cap = nreg;
            }
            reps[i].k.refstream = 1;
            reps[i].k.rstate = h->d_rstate;
        }
        reps[i].decide = nullptr;
        if (reps[i].k.act && !reps[i].k.roles && reps[i].k.lc == 0) split = true;
        if (reps[i].k.K != reps[0].k.K || reps[i].k.S != reps[0].k.S)
            return fail(CPM_EINVAL, "batched replicas need the same colour schedule (K, S)");
        const long nsub = 1L << (reps[i].lh * b->d);
        const long w = (std::max(nsub, (long)b->sims[i]->n + 1) + gcpm::kBlock - 1) / gcpm::kBlock;
        need[i] = std::max(1L, (long)std::ceil(w * b->block_scale));
    }
    // Batch form: 0 per-phase launches, 1 one persistent cooperative launch
    // per chunk, 2 per-phase launches captured once and replayed as a hip
    // graph; -1 chooses.  The persistent form needs 2D, shared roles and none
    // of the reference stream, cell_sync, persistence or the split act path.
    bool pers_any = false;
    for (int i = 0; i < n; i++) pers_any = pers_any || reps[i].k.pers;
    const bool persistent_ok = b->d == 2 && !reps[0].k.refstream && !reps[0].k.roles && !a->cell_sync &&
                               !pers_any && !split;
    int form = b->form;
    FormTuner::Entry* run_tune = nullptr;  // run-level choice among the forms
    int run_ti = -1;
    hipEvent_t run_e0 = nullptr, run_e1 = nullptr;
    const long total_mcs = (a->iterations / a->inner_iterations) * a->inner_iterations;
    if (form == -1) {
        // Per-phase launches unless the autotuner finds another form faster
        // (graph replay measured no faster than direct per-phase launches on
        // MI355X: equal on 64 replicas of 128^2-1024^2, 20-40% slower on
        // small launch-bound batches; the persistent form wins on a few tiny
        // replicas).
        form = 0;
        if (b->autotune && !b->timing && total_mcs > 0) {
            // Whole runs alternate between the candidates (each timed per
            // MCS on the batch stream) until each has its samples; then the
            // fastest is kept for the workload.
            std::vector<int> cands = {0, 2};
            if (persistent_ok) cands.push_back(1);
            uint64_t key = (uint64_t)1 << 63 | (uint64_t)b->d | (uint64_t)n << 2 | (uint64_t)reps[0].k.lc << 20 |
                           (uint64_t)reps[0].k.roles << 24 | (uint64_t)a->updates_per_checkerboard_switch << 25 |
                           (uint64_t)persistent_ok << 33 | (uint64_t)std::min(total_mcs, 1L << 20) << 34;
            for (int i = 0; i < n; i++) key = key * 1000003u + (uint64_t)reps[i].lh;
            key = key * 1000003u + (uint64_t)a->inner_iterations;
            run_tune = &b->tuner.entry(key, cands);
            if (run_tune->best >= 0) {
                form = run_tune->best;
                run_tune = nullptr;
            } else if ((run_ti = b->tuner.pick(*run_tune)) < 0) {
                run_tune = nullptr;
            } else {
                form = run_tune->forms[run_ti];
                HIP_TRY(b->tuner.events(&run_e0, &run_e1));
            }
        }
    }
    // A run-level sample counts when the run used its candidate without
    // capturing a graph (a capture's host time would be charged to it).
    auto finish_run_tune = [&](int used_form, int64_t captures0) -> int {
        if (!run_tune) return CPM_OK;
        HIP_TRY(hipEventRecord(run_e1, b->stream));
        if (used_form == run_tune->forms[run_ti] && b->graph_captures == captures0) {
            FormTuner::Pending pd{run_tune, run_ti, run_e0, run_e1};
            pd.scale = (double)total_mcs;
            b->tuner.pending.push_back(pd);
        }
        HIP_TRY(b->tuner.settle());
        return CPM_OK;
    };
    if (split) {
        // Act with shared roles and c = 1 in any replica: every replica's
        // phases are decided, then applied (for a replica without the term
        // the decisions are those of one launch: with shared roles and c = 1
        // a target's neighbourhood holds no other target).
        if (form == 1)
            return fail(CPM_ENOTIMPL,
                        "the act term with shared roles and positions_per_checkerboard = 1 needs the "
                        "per-phase batch form");
        for (int i = 0; i < n; i++) {
            cpm_handle h = b->sims[i];
            const size_t nsub = (size_t)1 << (reps[i].lh * b->d);
            if (nsub > h->decide_cap) {
                (void)hipFree(h->d_decide);
                h->d_decide = nullptr;
                HIP_TRY(hipMalloc(&h->d_decide, nsub));
                h->decide_cap = nsub;
            }
            reps[i].decide = h->d_decide;
        }
    }
    const long chunks = a->iterations / a->inner_iterations;
    if (chunks == 0) return CPM_OK;
    if (form != 1) {
        const int64_t captures0 = b->graph_captures;
        if (run_tune) HIP_TRY(hipEventRecord(run_e0, b->stream));
        if (int rc = batch_run_phase(b, a, reps, total_mcs, split, form == 2)) return rc;
        return finish_run_tune(b->last_form, captures0);
    }
    if (reps[0].k.refstream) return fail(CPM_ENOTIMPL, "the reference stream needs the per-phase batch form");
    if (reps[0].k.roles)
        return fail(CPM_ENOTIMPL, "the persistent batch form implements shared roles only (shared = 1)");
    if (a->cell_sync) return fail(CPM_ENOTIMPL, "cell_sync needs the per-phase batch form");
    for (int i = 0; i < n; i++)
        if (reps[i].k.pers) return fail(CPM_ENOTIMPL, "persistence needs the per-phase batch form");
    if (b->d == 3) return fail(CPM_ENOTIMPL, "the persistent batch form is off in 3D");
    for (int i = 0; i < n; i++) {  // it reads 32-bit sites
        if (int rc = b->sims[i]->ensure_wide()) return rc;
        b->sims[i]->narrow_valid = false;
    }

    // Block groups: each replica gets blocks in proportion to its work, at
    // least one, within the co-resident limit of the cooperative grid.
    int maxb = 0;
    HIP_TRY(gcpm::batch_max_blocks(b->d, &maxb));
    if (maxb < n) return fail(CPM_EINVAL, "more replicas than co-resident blocks");
    long want = 0;
    for (long v : need) want += v;
    std::vector<int> nb(n);
    int total = 0;
    for (int i = 0; i < n; i++) {
        long v = need[i];
        if (want > maxb) v = std::max(1L, (long)((double)need[i] * (maxb - n) / (double)want) + 1);
        nb[i] = (int)v;
        total += nb[i];
    }
    while (total > maxb) {  // trim rounding excess from the largest groups
        int j = (int)(std::max_element(nb.begin(), nb.end()) - nb.begin());
        nb[j]--;
        total--;
    }
    std::vector<int> rep_of(total);
    for (int i = 0, b0 = 0; i < n; i++) {
        reps[i].block0 = b0;
        reps[i].nblocks = nb[i];
        for (int j = 0; j < nb[i]; j++) rep_of[b0 + j] = i;
        b0 += nb[i];
    }
    if (total > b->rep_of_cap) {
        (void)hipFree(b->d_rep_of);
        b->d_rep_of = nullptr;
        HIP_TRY(hipMalloc(&b->d_rep_of, sizeof(int) * total));
        b->rep_of_cap = total;
    }
    b->last_blocks = total;
    // The batch stream waits for each replica's queued work.
    for (int i = 0; i < n; i++) {
        HIP_TRY(hipEventRecord(b->ready[i], b->sims[i]->stream));
        HIP_TRY(hipStreamWaitEvent(b->stream, b->ready[i], 0));
    }
    HIP_TRY(hipMemcpyAsync(b->d_rep_of, rep_of.data(), sizeof(int) * total, hipMemcpyHostToDevice, b->stream));
    gcpm::BatchArgs ba{};
    ba.reps = b->d_reps;
    ba.rep_of = b->d_rep_of;
    ba.bars = b->d_bars;
    ba.err = b->d_bars + 2 * n;
    ba.nmcs = a->inner_iterations;
    for (long ch = 0; ch < chunks; ch++) {
        for (int i = 0; i < n; i++)
            reps[i].mcs0 = (uint32_t)(b->sims[i]->mcs + (uint64_t)ch * a->inner_iterations);
        // Wait for the copy (and so the previous chunk) before reps is
        // rewritten for the next chunk.
        HIP_TRY(hipMemcpyAsync(b->d_reps, reps.data(), sizeof(gcpm::ReplicaDesc) * n, hipMemcpyHostToDevice,
                               b->stream));
        HIP_TRY(hipStreamSynchronize(b->stream));
        if (run_tune && ch == 0) HIP_TRY(hipEventRecord(run_e0, b->stream));
        HIP_TRY(gcpm::launch_batch(b->d, ba, total, b->stream));
    }
    b->last_form = 1;
    if (int rc = finish_run_tune(1, b->graph_captures)) return rc;
    if (int rc = check_barrier(b->d_bars, 2 * n + 1, b->stream)) return rc;
    // Each replica's later work waits for the batch.
    HIP_TRY(hipEventRecord(b->done, b->stream));
    for (int i = 0; i < n; i++) {
        HIP_TRY(hipStreamWaitEvent(b->sims[i]->stream, b->done, 0));
        b->sims[i]->mcs += (uint64_t)(chunks * a->inner_iterations);
    }
    return CPM_OK;
}

int cpm_batch_set_autotune(cpm_batch_handle b, int on) {
    if (!b) return fail(CPM_EINVAL, "null batch");
    b->autotune = on != 0;
    return CPM_OK;
}

int cpm_batch_set_kernel_options(cpm_batch_handle b, int small, int compact, int attempts) {
    if (!b) return fail(CPM_EINVAL, "null batch");
    for (int v : {small, compact})
        if (v < -1 || v > 1) return fail(CPM_EINVAL, "kernel options must be -1 (default), 0 or 1");
    if (attempts != -1 && attempts != 1 && attempts != 4 && attempts != 8)
        return fail(CPM_EINVAL, "batched attempts per thread must be -1 (default), 1, 4 or 8");
    b->opts.small = small;
    b->opts.compact = compact;
    b->opts.att = attempts;
    return CPM_OK;
}

int cpm_batch_set_form(cpm_batch_handle b, int form) {
    if (!b) return fail(CPM_EINVAL, "null batch");
    if (form < -1 || form > 2)
        return fail(CPM_EINVAL, "form must be -1 (choose), 0 (per phase), 1 (persistent) or 2 (graph replay)");
    b->form = form;
    return CPM_OK;
}

int cpm_batch_set_blocks(cpm_batch_handle b, double blocks_per_work) {
    if (!b) return fail(CPM_EINVAL, "null batch");
    if (!(blocks_per_work > 0.0) || blocks_per_work > 16.0)
        return fail(CPM_EINVAL, "blocks_per_work must be in (0, 16]");
    b->block_scale = blocks_per_work;
    return CPM_OK;
}

int cpm_batch_sync(cpm_batch_handle b) {
    if (!b) return fail(CPM_EINVAL, "null batch");
    HIP_TRY(hipStreamSynchronize(b->stream));
    return CPM_OK;
}

int cpm_batch_timing_begin(cpm_batch_handle b) {
    if (!b) return fail(CPM_EINVAL, "null batch");
    b->timing = true;
    b->ev_used = 0;
    b->t_launches = 0;
    b->t_updates = 0;
    return CPM_OK;
}

int cpm_batch_timing_end(cpm_batch_handle b, double* kernel_ms, int64_t* launches, int64_t* site_updates) {
    if (!b) return fail(CPM_EINVAL, "null batch");
    HIP_TRY(hipStreamSynchronize(b->stream));
    double total = 0.0;
    for (size_t i = 0; i + 1 < b->ev_used; i += 2) {
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, b->ev[i], b->ev[i + 1]));
        total += ms;
    }
    if (kernel_ms) *kernel_ms = total;
    if (launches) *launches = b->t_launches;
    if (site_updates) *site_updates = b->t_updates;
    b->timing = false;
    b->ev_used = 0;
    return CPM_OK;
}

const char* cpm_batch_last_kernel(cpm_batch_handle b) { return b ? b->last_kernel : ""; }

int cpm_batch_graph_info(cpm_batch_handle b, int* last_form, int64_t* captures, int64_t* replays) {
    if (!b) return fail(CPM_EINVAL, "null batch");
    if (last_form) *last_form = b->last_form;
    if (captures) *captures = b->graph_captures;
    if (replays) *replays = b->graph_replays;
    return CPM_OK;
}

int cpm_batch_info(cpm_batch_handle b, int* replicas, int* blocks) {
    if (!b) return fail(CPM_EINVAL, "null batch");
    if (replicas) *replicas = (int)b->sims.size();
    if (blocks) *blocks = b->last_blocks;
    return CPM_OK;
}

int cpm_energy_stats(cpm_handle h, int64_t* pairs, int64_t* moments) {
    return cpm_energy_stats_many(&h, 1, pairs, moments);
}

int cpm_energy_grad(cpm_handle h, int ntypes, const double* theta, double* H, double* dtheta) {
    if (int rc = check_handle(h)) return rc;
    if (ntypes < 1 || ntypes > CPM_NTYPES) return fail(CPM_EINVAL, "ntypes must be in [1, 20]");
    if (!theta || !H || !dtheta) return fail(CPM_EINVAL, "null argument");
    const int np = CPM_NTYPES * CPM_NTYPES, nm = CPM_NTYPES * 5;
    std::vector<int64_t> pairs(np), mom(nm);
    if (int rc = cpm_energy_stats_many(&h, 1, pairs.data(), mom.data())) return rc;
    // H = sum_{a<=b} J_ab N_ab + sum_t lam_A (sum (A - A0)^2) + lam_P (sum (P - P0)^2),
    // every term a polynomial in theta over the exact integer statistics.
    const int nt = ntypes, nj = nt * (nt + 1) / 2;
    double e = 0.0;
    int j = 0;
    for (int a = 0; a < nt; a++)
        for (int b = a; b < nt; b++, j++) {
            const double N = (double)pairs[a * CPM_NTYPES + b];
            e += theta[j] * N;
            dtheta[j] = N;
        }
    for (int t = 0; t < nt; t++) {
        const double n = (double)mom[5 * t], SA = (double)mom[5 * t + 1], SA2 = (double)mom[5 * t + 2];
        const double SP = (double)mom[5 * t + 3], SP2 = (double)mom[5 * t + 4];
        const double la = theta[nj + t], a0 = theta[nj + nt + t];
        const double lp = theta[nj + 2 * nt + t], p0 = theta[nj + 3 * nt + t];
        const double qa = SA2 - 2.0 * a0 * SA + n * a0 * a0;  // sum (A - A0)^2
        const double qp = SP2 - 2.0 * p0 * SP + n * p0 * p0;
        e += la * qa + lp * qp;
        dtheta[nj + t] = qa;
        dtheta[nj + nt + t] = la * (2.0 * n * a0 - 2.0 * SA);
        dtheta[nj + 2 * nt + t] = qp;
        dtheta[nj + 3 * nt + t] = lp * (2.0 * n * p0 - 2.0 * SP);
    }
    *H = e;
    return CPM_OK;
}

int cpm_energy_stats_many(const cpm_handle* sims, int n, int64_t* pairs, int64_t* moments) {
    if (!sims || n < 0) return fail(CPM_EINVAL, "null engine list");
    if (!pairs || !moments) return fail(CPM_EINVAL, "null output");
    constexpr int np = CPM_NTYPES * CPM_NTYPES, nm = CPM_NTYPES * 5, nw = np + nm;
    for (int i = 0; i < n; i++) {
        if (int rc = check_handle(sims[i])) return rc;
        if (!sims[i]->pushed) return fail(CPM_ESTATE, "energy statistics exist only after push_to_gpu");
    }
    if (n == 0) return CPM_OK;
    StatsRunner& S = stats_runner();
    std::lock_guard<std::mutex> guard(S.mu);
    if (int rc = S.ensure(n)) return rc;
    double ms_total = 0.0;
    for (int grp = 0; grp < 4; grp++) {
        // one launch per dimensionality and site width over its engines, each
        // read from whichever lattice copy is current (no conversion)
        const int d = 2 + (grp >> 1);
        const bool narrow_grp = grp & 1;
        std::vector<int> idx;
        for (int i = 0; i < n; i++)
            if (sims[i]->d == d && !sims[i]->wide_valid == narrow_grp) idx.push_back(i);
        if (idx.empty()) continue;
        std::vector<gcpm::StatsDesc> desc(idx.size());
        std::vector<int> eng_of;
        for (size_t j = 0; j < idx.size(); j++) {
            cpm_handle h = sims[idx[j]];
            const bool narrow = !h->wide_valid;
            gcpm::StatsDesc& e = desc[j];
            e.lat = narrow ? (const void*)h->d_lat16 : (const void*)h->d_lat;
            e.narrow = narrow ? 1 : 0;
            e.lgL = h->lgL;
            e.idbits = narrow ? (uint32_t)h->lat16_bits : 24u;
            e.nt = h->nt;
            e.tab = h->d_tab[0];
            e.ctype = h->d_ctype;
            e.ncells = h->n;
            e.out = S.d_out + (size_t)idx[j] * nw;
            e.pair0 = (int)eng_of.size();
            eng_of.insert(eng_of.end(), gcpm::stats_pair_blocks(d, h->lgL), (int)j);
            e.cell0 = (int)eng_of.size();
            eng_of.insert(eng_of.end(), gcpm::stats_cell_blocks(h->n), (int)j);
            // the statistics read the lattice after the engine's queued work
            HIP_TRY(hipEventRecord(S.ready, h->stream));
            HIP_TRY(hipStreamWaitEvent(S.stream, S.ready, 0));
        }
        if (int rc = S.ensure_map(eng_of.size())) return rc;
        HIP_TRY(hipMemsetAsync(S.d_out, 0, sizeof(unsigned long long) * nw * (size_t)n, S.stream));
        HIP_TRY(hipMemcpyAsync(S.d_desc, desc.data(), sizeof(gcpm::StatsDesc) * desc.size(),
                               hipMemcpyHostToDevice, S.stream));
        HIP_TRY(hipMemcpyAsync(S.d_map, eng_of.data(), sizeof(int) * eng_of.size(), hipMemcpyHostToDevice,
                               S.stream));
        bool t3 = narrow_grp && d == 2;  // the three-type chunk form applies to every engine
        for (const gcpm::StatsDesc& e : desc) t3 = t3 && e.nt <= 3 && e.lgL >= 3;
        HIP_TRY(gcpm::launch_stats_batch(d, narrow_grp, t3, S.d_desc, S.d_map, (int)eng_of.size(), S.stream,
                                          S.e0, S.e1));
        HIP_TRY(hipMemcpyAsync(S.h_out, S.d_out, sizeof(unsigned long long) * nw * (size_t)n,
                               hipMemcpyDeviceToHost, S.stream));
        HIP_TRY(hipStreamSynchronize(S.stream));  // (desc and eng_of are host temporaries)
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, S.e0, S.e1));
        ms_total += ms;
        for (int i : idx) {
            for (int j = 0; j < np; j++) pairs[(size_t)i * np + j] = (int64_t)S.h_out[(size_t)i * nw + j];
            for (int j = 0; j < nm; j++) moments[(size_t)i * nm + j] = (int64_t)S.h_out[(size_t)i * nw + np + j];
        }
    }
    S.last_ms = ms_total;
    return CPM_OK;
}

int cpm_energy_stats_many_noalloc(const cpm_handle* sims, int n, int64_t* pairs, int64_t* moments) {
    if (!sims || n < 0) return fail(CPM_EINVAL, "null engine list");
    if (!pairs || !moments) return fail(CPM_EINVAL, "null output");
    if (n > gcpm::kStatsScratchEngines)
        return fail(CPM_EINVAL, "cpm_energy_stats_many_noalloc: at most " +
                                    std::to_string(gcpm::kStatsScratchEngines) + " engines per call");
    for (int i = 0; i < n; i++) {
        if (int rc = check_handle(sims[i])) return rc;
        if (!sims[i]->pushed) return fail(CPM_ESTATE, "energy statistics exist only after push_to_gpu");
    }
    if (n == 0) return CPM_OK;
    constexpr int nw = gcpm::kStatsWords, np = CPM_NTYPES * CPM_NTYPES, nm = CPM_NTYPES * 5;
    unsigned long long* scr = nullptr;
    HIP_TRY(gcpm::stats_scratch(&scr));  // the device's own copy of the scratch array
    // Everything runs on the first engine's stream, after every engine's
    // queued work.
    hipStream_t s = sims[0]->stream;
    for (int i = 1; i < n; i++) {
        HIP_TRY(hipEventRecord(sims[i]->ev_ready, sims[i]->stream));
        HIP_TRY(hipStreamWaitEvent(s, sims[i]->ev_ready, 0));
    }
    HIP_TRY(gcpm::launch_fill_zero(scr, (size_t)n * nw, s));
    for (int grp = 0; grp < 4; grp++) {
        const int d = 2 + (grp >> 1);
        const bool narrow_grp = grp & 1;
        gcpm::StatsArgs a;  // (on the stack: a kernel argument)
        a.n = 0;
        int blocks = 0;
        for (int i = 0; i <= n; i++) {
            const bool take = i < n && sims[i]->d == d && !sims[i]->wide_valid == narrow_grp;
            if (take) {
                cpm_handle h = sims[i];
                gcpm::StatsDesc& e = a.d[a.n];
                e.lat = narrow_grp ? (const void*)h->d_lat16 : (const void*)h->d_lat;
                e.narrow = narrow_grp ? 1 : 0;
                e.lgL = h->lgL;
                e.idbits = narrow_grp ? (uint32_t)h->lat16_bits : 24u;
                e.nt = h->nt;
                e.tab = h->d_tab[0];
                e.ctype = h->d_ctype;
                e.ncells = h->n;
                e.out = scr + (size_t)i * nw;
                e.pair0 = blocks;
                blocks += gcpm::stats_pair_blocks(d, h->lgL);
                e.cell0 = blocks;
                blocks += gcpm::stats_cell_blocks(h->n);
                a.end[a.n] = blocks;
                a.n++;
            }
            if (a.n > 0 && (a.n == gcpm::kStatsArgEngines || i == n)) {
                bool t3 = narrow_grp && d == 2;  // (as in cpm_energy_stats_many)
                for (int j = 0; j < a.n; j++) t3 = t3 && a.d[j].nt <= 3 && a.d[j].lgL >= 3;
                HIP_TRY(gcpm::launch_stats_args(d, narrow_grp, t3, a, blocks, s));
                a.n = 0;
                blocks = 0;
            }
        }
    }
    HIP_TRY(hipMemcpy2DAsync(pairs, np * sizeof(int64_t), scr, nw * sizeof(unsigned long long),
                             np * sizeof(int64_t), n, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpy2DAsync(moments, nm * sizeof(int64_t), scr + np, nw * sizeof(unsigned long long),
                             nm * sizeof(int64_t), n, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return CPM_OK;
}

int cpm_energy_stats_timing(double* kernel_ms) {
    if (!kernel_ms) return fail(CPM_EINVAL, "null output");
    StatsRunner& S = stats_runner();
    std::lock_guard<std::mutex> guard(S.mu);
    *kernel_ms = S.last_ms;
    return CPM_OK;
}

}  // extern "C"
