// batch.h -- many independent umcx simulations in one persistent GPU launch.
//
// Every job keeps the semantics of a separate MCX_run_simulation (src/umcx.cpp:783-848): its
// own label volume, media, source, time gate, seed words and photon count. Photon i of job j is
// the same photon a single-job run would trace, so per-job results equal single-job results up
// to float summation order.
//
// Scheduling: a workgroup works on one job at a time. Its waves take photons of that job from
// the job's device cursor (work stealing at photon granularity, one atomic per wave refill) and
// tally deposits that fall in the job's LDS tile in LDS. When the job runs dry the workgroup
// flushes the tile to HBM and moves to the job with the most estimated work left per workgroup
// already on it. The tally-zeroing kernels and the launch are captured once in a hipGraph and
// replayed per run.
//
// Tuning (at creation): a counting pilot of every job gives each job's deposit map, hence its
// cost per photon and its LDS tile for every candidate tile size; short trial launches over a
// 1/32 sample of every job then pick the workgroup size and workgroups per CU (so the LDS share
// and the tile size). Tiled shapes are preferred over untiled ones for the accuracy of the
// float32 tallies; tally_mode 0 switches tiles off.
#pragma once
#include <cstdint>
#include <vector>

#include "config.h"
#include "runtime.h"
#include "tally_box.h"

namespace umcx {

struct BatchTrial {
    int block = 0, blocks = 0;   // workgroup size, persistent workgroups
    int tile_cells = 0;          // LDS tally cells per workgroup (0: global atomics only)
    double ms = 0.0;             // trial launch time
    double tile_share = 0.0;     // pilot deposits inside the tiles, over all jobs
};

class BatchSession {
public:
    BatchSession(const std::vector<Problem>& jobs, const LaunchConfig& lc);
    ~BatchSession();
    BatchSession(const BatchSession&) = delete;
    BatchSession& operator=(const BatchSession&) = delete;

    RunStats run();                                   // zero the tallies, run every job's photons
    int njobs() const { return (int)probs_.size(); }
    const Problem& problem(int j) const { return probs_.at(j); }
    uint64_t fluence_len(int j) const;
    void fetch_fluence(int j, float* out) const;      // raw tallies [t][x][y][z]
    void energies(int j, double out[3]) const;        // escaped, launched, deposited
    // detected photons of job j (:155-184): all detections and rows kept (<= its MaxDetPhoton),
    // rows [saved x detp_columns] plus the photon index and final weight of each (any may be null)
    uint32_t detected(int j) const;
    uint32_t saved(int j) const;
    void fetch_detp(int j, float* rows, uint32_t* ids, float* weights) const;
    int block() const { return chosen_.block; }
    int blocks() const { return chosen_.blocks; }
    int tile_cells() const { return chosen_.tile_cells; }
    const TallyBox& box(int j) const { return boxes_.at(j); }
    const std::vector<BatchTrial>& trials() const { return trials_; }
    double cost(int j) const { return cost_.at(j); }   // pilot deposits per photon

private:
    std::vector<Problem> probs_;
    LaunchConfig lc_;
    BatchTrial chosen_;
    std::vector<BatchTrial> trials_;
    std::vector<TallyBox> boxes_;
    std::vector<double> cost_;
    std::vector<uint64_t> flu_off_;
    uint64_t flu_total_ = 0;
    int cus_ = 0;
    void* stream_ = nullptr;
    void* ev0_ = nullptr;
    void* ev1_ = nullptr;
    void* graph_ = nullptr;
    void* graph_exec_ = nullptr;
    void* d_labels_ = nullptr;
    void* d_media_ = nullptr;
    void* d_jobs_ = nullptr;
    int32_t* d_first_ = nullptr;          // initial job of each workgroup
    unsigned long long* d_jcur_ = nullptr;  // photon cursor per job
    uint32_t* d_active_ = nullptr;        // workgroups on each job
    float* d_fluence_ = nullptr;
    double* d_energy_ = nullptr;
    int first_cap_ = 0;
    // detected photons: per job, its detectors' offset and count in the detector table and its
    // row slots (the first slot's index in the id and weight buffers, its float offset in the
    // row buffer)
    struct JobRows {
        uint32_t det_off = 0, detnum = 0, maxdet = 0, detlen = 0, ppathlen = 0;
        uint64_t row_off = 0, detp_off = 0;
    };
    std::vector<JobRows> rows_;
    uint64_t rows_total_ = 0, detp_floats_ = 0;
    uint32_t pp_max_ = 0;
    bool any_det_ = false;
    uint32_t* d_ndet_ = nullptr;
    float* d_detpos_ = nullptr;
    float* d_detp_ = nullptr;
    uint32_t* d_detid_ = nullptr;
    float* d_detw_ = nullptr;
    size_t lds_fixed(int block) const;
    void set_rows(void* args) const;

    // upload a job table (nphoton scaled down to `sample` photons when nonzero, tiles from boxes)
    // and the workgroups' initial jobs, then launch once on the session stream
    double launch_once(int block, int blocks, int tile_cells, const std::vector<TallyBox>* boxes,
                       uint64_t sample_div, uint64_t sample_min, uint64_t sample_max, int count_mode);
    void upload_jobs(const std::vector<TallyBox>* boxes, uint64_t sample_div, uint64_t sample_min,
                     uint64_t sample_max, std::vector<double>& work);
    void upload_first(int blocks, const std::vector<double>& work);
    void zero_state(bool fluence);
    void tune();
};

}  // namespace umcx
