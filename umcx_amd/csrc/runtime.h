// runtime.h -- device-side state of one simulation and the launches that fill it.
//
// A Session uploads one Problem to the GPU once (labels as uint8 when the media fit in 7 bits,
// media and detectors, zeroed tallies) and then runs photon ranges [first, first+count) of
// the MCX_kernel loop (src/umcx.cpp:752-775) with the hand-written gfx950 kernel. Tallies
// accumulate across runs until reset(), so a photon range can be split over calls.
#pragma once
#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "config.h"

namespace umcx {

struct LaunchConfig {
    int device = 0;          // HIP device ordinal
    int block = 0;           // threads per workgroup (0: tuned default)
    int blocks = 0;          // persistent workgroups (0: fill the GPU)
    int tally_mode = -1;     // -1 tuned, 0 global atomics, 1 LDS-cached tile
    int vote = -1;           // scatter-round threshold in lanes (0: scatter right away, -1: tuned)
    int label_cache = 0;     // 0: labels in LDS when they fit, -1: labels read from HBM
    int deposit_cache = 0;   // 0: chosen from the pilot, 1: always, -1: never
    int deposit_log = 0;     // global deposits to a log summed by separate passes: 0 chosen from the
                             // pilot, 1 always (grids below 2^26 cells), -1 never
    int deposit_log_segments = 0;   // log capacity in segments of 4096 records (0: sized per run)
    bool trace = false;      // keep each photon's final state (parity tests; small runs only)
};

// Final state of one photon, for per-photon parity checks.
struct TraceRec {
    float pos[4], vec[4], len[4];
    int32_t mediaid, lastvoxelidx;
    float w0;
    uint32_t pad;
};

struct RunStats {
    double kernel_ms = 0.0;  // HIP-event time of the photon kernel(s) on the session stream
    uint64_t photons = 0;
};

class Session {
public:
    Session(const Problem& p, const LaunchConfig& lc);
    ~Session();
    Session(const Session&) = delete;
    Session& operator=(const Session&) = delete;

    void reset();                                        // zero fluence, detections, energies
    RunStats run(uint64_t first, uint64_t count);        // launch and wait
    // host copies of the raw (unnormalised) tallies
    void fetch_fluence(float* out) const;                // maxgate*nx*ny*nz, [t][x][y][z]
    uint32_t detected() const;                           // all detections, incl. dropped
    uint32_t saved() const;                              // rows kept (<= maxdetphotons)
    void fetch_detp(float* rows, uint32_t* ids, float* weights) const;
    void energies(double out[3]) const;                  // escaped, launched, deposited (fp64)
    // per-voxel mua/mus in 1/mm overriding the labels' (null: back to labels); device or host
    void set_voxel_props(const float* mua, const float* mus, bool on_device);
    // move the source (position, direction) for the next runs
    void set_source(const float pos[3], const float dir[3]);
    // several sources in one run (host table {pos xyz, dir xyz} per source; null: off): photon
    // slot p runs photon p mod per_source of source p / per_source, seeded as in its own run
    void set_sources(const float* table, int nsrc, uint64_t per_source);
    // re-trace photons ids[0..n) (device arrays) adding coef[k] x perturbation-MC derivatives of
    // the photon's weight (mua) and path density (mus) into slice slices[k] of gmua / gmus
    // (device, slice stride `stride` floats, voxels in label order); final weights to final_w.
    // With `sources` (device, {pos xyz, dir xyz} per source), record k is launched from source
    // src_index[k] instead of the session's source.
    RunStats replay(const uint32_t* ids, const int32_t* slices, const float* coef, uint64_t n, float* gmua,
                    float* gmus, uint64_t stride, float* final_w, const int32_t* src_index = nullptr,
                    const float* sources = nullptr);
    uint32_t* device_detid() const { return d_detid_; }
    float* device_detw() const { return d_detw_; }
    float* device_detp() const { return d_detp_; }
    void fetch_trace(TraceRec* out, uint64_t n) const;   // of the last run
    uint64_t fluence_len() const { return fluence_len_; }
    const Problem& problem() const { return prob_; }
    int block() const { return block_; }
    int blocks() const { return blocks_; }
    // shape of the last run() launch (half the lanes for short runs, see run_shape)
    int last_block() const { return last_block_; }
    int last_blocks() const { return last_blocks_; }
    int tally_mode() const { return tile_[4] > 0 ? 1 : 0; }   // effective: 1 when an LDS tile is in use
    int lds_label_bits() const { return plabel_words_ ? (int)plabel_bits_ : 0; }   // 0: labels read from HBM
    // the cache keys are 32-bit fluence indices: grids of 2^32 - 1 cells or more never use it
    bool deposit_cache() const {
        return fluence_len_ < 0xFFFFFFFFull && (lc_.deposit_cache > 0 || (lc_.deposit_cache == 0 && dep_cache_));
    }
    double deposit_cache_hits() const { return dep_hits_; }   // from the last pilot (0 before one)
    // global deposits go to the deposit log (summed into the grid by separate passes) instead of
    // float atomics: chosen by the pilot when they would saturate the memory-side atomics
    bool deposit_log() const;
    double global_deposits_per_photon() const { return dlog_rpp_; }   // pilot estimate (0 before one)
    int vote() const { return lc_.vote < 0 ? vote_auto_ : std::min(lc_.vote, 64); }   // scatter-round threshold
    void* stream() const { return stream_; }
    // device pointers, for callers that keep results on the GPU
    float* device_fluence() const { return d_fluence_; }

private:
    Problem prob_;
    LaunchConfig lc_;
    int block_ = 256, blocks_ = 0, tally_mode_ = 0;
    int cu_count_ = 1, last_block_ = 0, last_blocks_ = 0;
    bool auto_shape_ = false;            // launch shape chosen per run (no explicit block, blocks, threads)
    int label_bytes_ = 1;
    uint64_t fluence_len_ = 0;
    void* stream_ = nullptr;
    void* ev0_ = nullptr;
    void* ev1_ = nullptr;
    void* d_labels_ = nullptr;
    float* d_media_ = nullptr;
    float* d_detpos_ = nullptr;
    float* d_fluence_ = nullptr;
    float* d_detp_ = nullptr;
    uint32_t* d_detid_ = nullptr;
    float* d_detw_ = nullptr;
    uint32_t* d_counters_ = nullptr;     // [0] detections
    unsigned long long* d_cursor_ = nullptr;
    double* d_energy_ = nullptr;
    TraceRec* d_trace_ = nullptr;
    void* d_vprops_ = nullptr;
    uint64_t trace_cap_ = 0;
    uint64_t last_count_ = 0;
    uint32_t detcap_ = 0;
    // LDS tally tile {x0, y0, z0, g0, sx, sy, sz, sg}; sx = 0 means none
    int tile_[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    bool tile_chosen_ = false;
    bool dep_cache_ = true;              // per-lane deposit cache on (off when the tile takes most deposits)
    double dep_hits_ = 0.0;              // share of the pilot's global deposits that hit the deposit cache
    uint64_t tile_budget_ = 0;
    float* d_msrc_ = nullptr;            // multi-source table {pos xyz, dir xyz} per source
    uint64_t msrc_n_ = 0;                // photons per source (0: one source)
    uint64_t msrc_count_ = 0;            // sources in the table
    uint32_t* d_plabels_ = nullptr;      // packed media ids for LDS staging (null: not staged)
    int label_mode() const { return plabel_words_ ? 1 : 0; }   // kernels' label mode: staged ids or bytes
    uint32_t plabel_words_ = 0, plabel_bits_ = 0, plabel_id0_ = 0, plabel_step_ = 1;
    // the 1-bit form of the staged labels, kept beside the 2-/4-bit one until the pilot picks
    uint32_t* d_plabels1_ = nullptr;
    uint32_t plabel1_words_ = 0, plabel1_id0_ = 0, plabel1_step_ = 0;
    size_t per_wg_lds_ = 0;     // LDS per workgroup at the session's launch shape
    // deposit log: records {cell, value}, the same records grouped by bucket, each segment's
    // fill, the segment counter, the (bucket, workgroup) counts and the bucket starts
    int vote_auto_ = 0;                  // scatter-round lane vote chosen by the pilot (0: none)
    bool dlog_auto_ = false;             // the pilot chose the log
    bool dlog_unavailable_ = false;      // the device could not hold a log: float atomics
    double dlog_rpp_ = 0.0;              // global deposits per photon expected from the pilot
    void* d_dlog_ = nullptr;
    void* d_dlog_sorted_ = nullptr;
    uint32_t* d_dlog_fill_ = nullptr;
    unsigned long long* d_dlog_next_ = nullptr;
    uint32_t* d_dlog_hist_ = nullptr;
    uint32_t* d_dlog_start_ = nullptr;
    uint64_t dlog_records_ = 0;          // capacity in records (a multiple of the segment size)
    void ensure_log(uint64_t count);
    bool log_ready(uint64_t count);
    void enqueue_log_passes(void* stream, uint64_t expect_records);

    void* build_args(uint64_t first, uint64_t count, void* out) const;
    size_t lds_bytes(uint32_t tile_n, bool paths_needed = true) const;   // dynamic LDS of a launch
    void choose_tile(uint64_t first, uint64_t count);
    void run_shape(uint64_t count, int* block, int* blocks) const;

public:
    // the chosen LDS tally tile, {x0, y0, z0, g0, sx, sy, sz, sg} (all zero when none)
    const int* tile() const { return tile_; }
};

// Replica checks: evaluate the device libm replicas on n inputs (fn: 0 logf, 1 sinf, 2 cosf,
// 3 acosf, 4 expf) on the GPU.
void device_libm_eval(int fn, const float* in, float* out, uint64_t n, int device);
// Self-check of the short correctly rounded sqrt, reciprocal and division sequences against the
// full ones (which: 0 sqrt, 1 reciprocal, 2 division); returns the number of differing inputs.
uint64_t device_arith_check(int which, uint64_t* tested, int device);

// The sessions' tally-box search (densest cell, total, greedy growth, all on the device) applied
// to a host count grid [t][x][y][z].
struct TallyBoxSearch {
    int box[8];          // {x0, y0, z0, g0, sx, sy, sz, sg}, zeros for an empty grid
    double in_box, total;
};
TallyBoxSearch device_choose_box(const float* counts, int nx, int ny, int nz, int nt, uint64_t budget, int device);

int device_count();
std::string device_name(int device);

}  // namespace umcx

struct umcx_device_problem;
namespace umcx {
// umcx_run_device_noalloc: the photon kernel on caller-owned device buffers, enqueued on
// `stream` (a hipStream_t) of the current device, with no allocation (photon_kernel.hip).
void run_device_noalloc(const umcx_device_problem* q, uint64_t first, uint64_t count, void* stream);

}  // namespace umcx
