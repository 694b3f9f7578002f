// batch.h -- many independent umcx simulations in one persistent GPU launch.
//
// Every job keeps the semantics of a separate MCX_run_simulation (src/umcx.cpp:783-848): its
// own label volume, media, source, time gate, seed words and photon count. Photons of all jobs
// are cut into fixed-size chunks that are interleaved round-robin across jobs; lanes take photon
// slots from one device cursor (work stealing at photon granularity), so small and large jobs
// share the GPU instead of running one after another. Photon i of job j is the same photon a
// single-job run would trace, so per-job results equal single-job results up to float summation
// order. The zero-fill of the tallies and the launch are captured once in a hipGraph and
// replayed per run.
#pragma once
#include <cstdint>
#include <vector>

#include "config.h"
#include "runtime.h"

namespace umcx {

class BatchSession {
public:
    BatchSession(const std::vector<Problem>& jobs, const LaunchConfig& lc, uint32_t chunk = 2048);
    ~BatchSession();
    BatchSession(const BatchSession&) = delete;
    BatchSession& operator=(const BatchSession&) = delete;

    RunStats run();                                   // zero the tallies, run every job's photons
    int njobs() const { return (int)probs_.size(); }
    const Problem& problem(int j) const { return probs_.at(j); }
    uint64_t fluence_len(int j) const;
    void fetch_fluence(int j, float* out) const;      // raw tallies [t][x][y][z]
    void energies(int j, double out[3]) const;        // escaped, launched, deposited
    int block() const { return block_; }
    int blocks() const { return blocks_; }
    uint64_t slots() const { return slots_; }

private:
    std::vector<Problem> probs_;
    LaunchConfig lc_;
    int block_ = 256, blocks_ = 0;
    uint32_t chunk_ = 2048;
    uint64_t slots_ = 0;
    size_t lds_ = 0;
    bool media_in_lds_ = true;
    std::vector<uint64_t> flu_off_;
    uint64_t flu_total_ = 0;
    void* stream_ = nullptr;
    void* ev0_ = nullptr;
    void* ev1_ = nullptr;
    void* graph_ = nullptr;
    void* graph_exec_ = nullptr;
    void* d_labels_ = nullptr;
    void* d_media_ = nullptr;
    void* d_jobs_ = nullptr;
    void* d_chunks_ = nullptr;
    float* d_fluence_ = nullptr;
    double* d_energy_ = nullptr;
    unsigned long long* d_cursor_ = nullptr;
    int total_media_ = 0;
};

}  // namespace umcx
