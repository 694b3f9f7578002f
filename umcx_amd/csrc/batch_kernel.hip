// batch_kernel.hip -- persistent gfx950 kernel for many independent simulations (batch.h).
//
// The per-photon transport is the plain kernel's (photon_device.h); what changes is where a
// workgroup gets its photons and its job's constants. A workgroup holds one job at a time: the
// job's constants are workgroup-uniform (scalar registers), its media sit in LDS, and deposits
// inside the job's tally box accumulate in an LDS tile that is flushed to HBM when the job runs
// dry. Each wave takes the job's photons from the job's cursor, so waves and lanes stay busy
// until the job has no photon left; then the workgroup moves to the job with the most estimated
// work left per workgroup already on it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>

#include "batch.h"
#include "photon_device.h"

#define HIP_OK(x)                                                                                   \
    do {                                                                                            \
        hipError_t e_ = (x);                                                                        \
        if (e_ != hipSuccess)                                                                       \
            throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_) + " at " + \
                                     __FILE__ + ":" + std::to_string(__LINE__));                   \
    } while (0)

namespace umcx {
namespace dev {

constexpr uint32_t kMaxBatchMedia = 128;   // labels travel as uint8 with the detector bit in bit 7
// a wave takes new photons once this many of its lanes need one (or all do): the refill's cursor
// atomic and launch code then serve two lanes at a time (256-job workload 9,977 -> 9,914 ms; four
// lanes: 10,037 ms)
constexpr int kBatchRefillMin = 2;

// one job's constants (MCX_param :95-99 + source + seeds + buffer offsets + LDS tally box)
struct __align__(16) JobDesc {
    float4 srcpos, srcdir, srcp1, srcp2;
    uint32_t nx, ny, nz, nyz;
    uint64_t label_off, flu_off;
    uint64_t nphoton, nxyz;
    uint32_t seeds[4];
    float tstart, tend, rtstep;
    int32_t maxgate;
    uint32_t media_off, mediumnum;
    int32_t srctype, isreflect, outputtype, issavevol;
    int32_t tx0, ty0, tz0, tg0;        // LDS tally box origin (voxel, gate)
    uint32_t tsx, tsy, tsz, tsg;       // and extent; tsx == 0: no box
    float cost;                        // estimated work per photon (scheduling weight, > 0)
    // detected photons (:155-184): detectors at det_off of the detector table; maxdet row slots
    // starting at slot row_off of the id and weight buffers and at float detp_off of the row
    // buffer, detlen floats each, columns from savedetflag
    uint32_t detnum, det_off, maxdet;
    uint32_t detlen;
    uint64_t row_off, detp_off;
    uint32_t ppathlen;
    int32_t savedetflag;
    uint32_t pad[2];
};

struct BatchArgs {
    const JobDesc* jobs;
    uint32_t njobs;
    uint32_t tile_cap;          // LDS tally cells per workgroup (0: global atomics only)
    const int32_t* first;       // initial job of each workgroup
    const uint8_t* labels;      // all jobs' labels, bit 7 = detector mask
    const float4* media;        // all jobs' media
    float* fluence;
    double* energy;             // 3 per job
    unsigned long long* jcur;   // photon cursor per job
    uint32_t* active;           // workgroups on each job
    int32_t count_mode;         // 1: tally the number of deposits (pilot)
    // detected photons (SAVEDET kernels): every job's detectors, the row buffers and one
    // detection counter per job
    const float4* detpos;
    float* detp;                // rows; job j's start at its detp_off
    uint32_t* detid;            // photon index of each row (within its job)
    float* detw;                // final weight of each row
    uint32_t* ndet;             // detections per job, including ones beyond its maxdet
    uint32_t pp_max;            // partial-path columns per lane in LDS (the largest ppathlen)
};

__device__ __forceinline__ int box_local(int ix, int iy, int iz, int x0, int y0, int z0, uint32_t sx, uint32_t sy,
                                         uint32_t sz) {
    const unsigned dx = (unsigned)(ix - x0), dy = (unsigned)(iy - y0), dz = (unsigned)(iz - z0);
    return (dx < sx && dy < sy && dz < sz) ? (int)((dx * sy + dy) * sz + dz) : -1;
}

// Called by one full wave: the job with the most estimated work left per workgroup on it
// (photons not yet taken x cost per photon / (workgroups on it + 1)), -1 when every job is dry.
__device__ int pick_job(const BatchArgs& b) {
    float best = 0.f;
    int bj = -1;
    for (uint32_t j = __lane_id(); j < b.njobs; j += 64) {
        const uint64_t n = b.jobs[j].nphoton;
        const uint64_t c = __hip_atomic_load(b.jcur + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (c < n) {
            const uint32_t act = __hip_atomic_load(b.active + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const float s = (float)(n - c) * b.jobs[j].cost / (float)(act + 1u);
            if (bj < 0 || s > best) {
                best = s;
                bj = (int)j;
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        const float ob = __shfl_xor(best, o);
        const int oj = __shfl_xor(bj, o);
        if (oj >= 0 && (bj < 0 || ob > best || (ob == best && oj < bj))) {
            best = ob;
            bj = oj;
        }
    }
    return bj;
}

// SAVEDET: jobs may save detected photons (their partial paths per medium live per lane in LDS,
// after the tally tile)
template <bool SAVEDET>
__global__ void __launch_bounds__(1024) batch_kernel(BatchArgs b) {
    // LDS: libm tables | the current job's media | tally tile | per-lane partial paths
    extern __shared__ float4 lds4[];
    __shared__ int s_job;
    libm::LogfEntry* slog = (libm::LogfEntry*)lds4;
    uint64_t* sexp = (uint64_t*)(lds4 + 16);
    float4* smedia = lds4 + 32;
    float* stile = (float*)(lds4 + 32 + kMaxBatchMedia);
    float* sppath = stile + b.tile_cap;
    const int tid = threadIdx.x, bd = blockDim.x;
    if (tid < 16) slog[tid] = libm::kLogfTab[tid];
    if (tid < 32) sexp[tid] = libm::kExp2Tab[tid];
    if (tid == 0) {
        const int j = b.first[blockIdx.x];
        atomicAdd(b.active + j, 1u);
        s_job = j;
    }
    __syncthreads();
    const Tabs tabs{slog, sexp};

    for (;;) {
        const int job = __builtin_amdgcn_readfirstlane(s_job);
        if (job < 0) break;
        const JobDesc* J = b.jobs + job;
        const Grid g{J->nx, J->ny, J->nz, J->nyz};
        const uint64_t lab_off = J->label_off, flu_off = J->flu_off, nxyz = J->nxyz, nphoton = J->nphoton;
        const float tstart = J->tstart, tend = J->tend, rtstep = J->rtstep;
        const int maxgate = J->maxgate;
        const bool isreflect = J->isreflect != 0, energy_out = J->outputtype == kEnergy, savevol = J->issavevol != 0;
        const int tx0 = J->tx0, ty0 = J->ty0, tz0 = J->tz0, tg0 = J->tg0;
        const uint32_t tsy = J->tsy, tsz = J->tsz, tsg = J->tsg;
        const uint32_t tcells = J->tsx * tsy * tsz;
        const uint32_t tile_n = (b.tile_cap && tcells * tsg <= b.tile_cap) ? tcells * tsg : 0u;
        const uint32_t tsx = tile_n ? J->tsx : 0u;
        const uint32_t nmed = J->mediumnum;
        const uint32_t ppathlen = SAVEDET ? J->ppathlen : 0u;
        for (uint32_t i = tid; i < nmed; i += bd) smedia[i] = b.media[J->media_off + i];
        for (uint32_t i = tid; i < tile_n; i += bd) stile[i] = 0.f;
        __syncthreads();

        Photon p;
        Rng r;
        uint64_t pid = 0;
        bool alive = false, dry = false;
        float w0 = 0.f;
        double e_esc = 0.0, e_tot = 0.0, e_dep = 0.0, ph_dep = 0.0;
        while (true) {
            // ---- lanes without a photon take the job's next ones (one atomic per wave) ----
            const unsigned long long idle = __ballot(!alive);
            if (dry) {
                if (idle == ~0ull) break;
            } else if (__popcll(idle) >= kBatchRefillMin || idle == ~0ull) {
                const int leader = __ffsll((long long)idle) - 1;
                unsigned long long base = 0;
                if ((int)__lane_id() == leader) base = atomicAdd(b.jcur + job, (unsigned long long)__popcll(idle));
                base = __shfl(base, leader);
                if (base + (unsigned long long)__popcll(idle) >= nphoton) dry = true;
                const unsigned long long k = base + (unsigned long long)__popcll(idle & lanemask_lt());
                if (!alive && k < nphoton) {
                    const uint32_t i32 = (uint32_t)k;
                    rng_seed(r, J->seeds[0] ^ i32, J->seeds[1] | i32, J->seeds[2] ^ i32, J->seeds[3] | i32);   // :761
                    const Source src{J->srcpos, J->srcdir, J->srcp1, J->srcp2, J->srctype};
                    w0 = launch(p, r, src);
                    ph_dep = 0.0;
                    pid = k;
                    if (SAVEDET)
                        for (uint32_t m = 0; m < ppathlen; m++) sppath[m * bd + tid] = 0.f;
                    p.lastvox = vox_index(g, p.ix, p.iy, p.iz);
                    bool entered = p.lastvox >= 0 || skip_to_box(p, g);
                    if (entered && p.lastvox < 0) entered = false;
                    if (entered) {
                        const uint32_t lab = b.labels[lab_off + (uint64_t)p.lastvox];
                        p.med = (lab & 0x7Fu) | ((lab & 0x80u) << 24);
                        const float4 m = smedia[p.med & kMed];
                        p.cmua = m.x;
                        p.cmus = m.y;
                        p.tvox = box_local(p.ix, p.iy, p.iz, tx0, ty0, tz0, tsx, tsy, tsz);
                        p.lx = rng_scatlen(r, tabs);
                        alive = true;
                    }
                }
            }
            if (!alive) continue;

            // ---- one step of sprint (:282-311) ----
            bool ended = false;
            if (p.lx > 0.f) {
                const float4 prop = smedia[p.med & kMed];
                uint32_t nlab;
                int nv = step(p, prop.w, g, tabs, GlobalLabels<uint8_t>{b.labels + lab_off}, nlab);
                if (nv != p.lastvox) {
                    const uint32_t mcur = p.med & kMed;
                    if (SAVEDET && ppathlen && mcur > 0) sppath[(mcur - 1) * bd + tid] += p.lz;   // :286-288
                    if (savevol) {
                        int tshift = (int)fminf((float)(maxgate - 1), (float)(int)floorf((p.ly - tstart) * rtstep));
                        float dw = p.lw - p.pw;
                        float val = energy_out ? dw : ((p.cmua < kEps) ? (p.lw * p.lz) : dw / p.cmua);
                        ph_dep += (double)dw;
                        if (b.count_mode) val = 1.f;
                        if (tshift >= 0) {
                            const unsigned gg = (unsigned)(tshift - tg0);
                            if (p.tvox >= 0 && gg < tsg)
                                __hip_atomic_fetch_add(stile + gg * tcells + (unsigned)p.tvox, val, __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_WORKGROUP);
                            else
                                unsafeAtomicAdd(b.fluence + flu_off + (uint64_t)p.lastvox + (uint64_t)tshift * nxyz, val);
                        }
                        p.lw = p.pw;
                        p.lz = 0.f;
                    }
                    if (p.ly > tend) {
                        ended = true;
                    } else {
                        uint32_t nm = (nv >= 0) ? GlobalLabels<uint8_t>{nullptr}.decode(nlab) : 0u;
                        const float4 nprop = smedia[nm & kMed];
                        if (isreflect && prop.w != nprop.w) {
                            if (boundary(p, prop.w, nprop.w, r, nv, nm) && (nv < 0 || (nm & kMed) == 0)) ended = true;
                        } else if (nv < 0 || ((nm & kMed) == 0 && mcur != 0)) {
                            ended = true;
                        }
                        if (!ended) {
                            const bool moved = nv != p.lastvox;
                            if (moved) p.tvox = box_local(p.ix, p.iy, p.iz, tx0, ty0, tz0, tsx, tsy, tsz);
                            p.lastvox = nv;
                            p.med = nm;
                            if (moved) {
                                p.cmua = nprop.x;
                                p.cmus = nprop.y;
                            }
                        }
                    }
                }
            }
            if (!ended) {
                if (!(p.lx > 0.f)) scatter(p, r, smedia[p.med & kMed], tabs);
                continue;
            }
            // detection (:276-278, :429-435, :164-176)
            if (SAVEDET && J->detnum && p.ly <= tend && (p.med & kDet)) {
                for (uint32_t d = 0; d < J->detnum; d++) {
                    const float4 dp = b.detpos[J->det_off + d];
                    const float dx = dp.x - p.px, dy = dp.y - p.py, dz = dp.z - p.pz;
                    if (dx * dx + dy * dy + dz * dz < dp.w * dp.w) {
                        const uint32_t slot = atomicAdd(b.ndet + job, 1u);
                        if (slot < J->maxdet) {
                            const uint64_t rowi = J->row_off + slot;
                            float* row = b.detp + J->detp_off + (uint64_t)slot * J->detlen;
                            uint32_t c = 0;
                            if (J->savedetflag & kDetID) row[c++] = (float)(d + 1);
                            for (uint32_t m = 0; m < ppathlen; m++) row[c++] = sppath[m * bd + tid];
                            if (J->savedetflag & kExitPos) { row[c++] = p.px; row[c++] = p.py; row[c++] = p.pz; }
                            if (J->savedetflag & kExitDir) { row[c++] = p.vx; row[c++] = p.vy; row[c++] = p.vz; }
                            b.detid[rowi] = (uint32_t)pid;
                            b.detw[rowi] = p.pw;
                        }
                    }
                }
            }
            if (p.med & kMed) {                                                                      // :765-770
                e_tot += (double)w0;
                e_esc += (double)p.pw;
                e_dep += ph_dep;
            }
            alive = false;
        }

        // ---- the job ran dry for this workgroup: energies, tile, next job ----
        e_esc = wave_sum(e_esc);
        e_tot = wave_sum(e_tot);
        e_dep = wave_sum(e_dep);
        if (__lane_id() == 0 && (e_tot != 0.0 || e_esc != 0.0 || e_dep != 0.0)) {
            unsafeAtomicAdd(b.energy + 3 * job + 0, e_esc);
            unsafeAtomicAdd(b.energy + 3 * job + 1, e_tot);
            unsafeAtomicAdd(b.energy + 3 * job + 2, e_dep);
        }
        __syncthreads();
        if (tile_n) {
            const uint32_t plane = tsy * tsz;
            for (uint32_t i = tid; i < tile_n; i += bd) {
                const float v = stile[i];
                if (v != 0.f) {
                    const uint32_t gg = i / tcells, c = i - gg * tcells;
                    const uint32_t dx = c / plane, rem = c - dx * plane, dy = rem / tsz, dz = rem - dy * tsz;
                    const uint64_t idx = flu_off + (uint64_t)(tg0 + (int)gg) * nxyz + (uint64_t)(tx0 + (int)dx) * g.nyz +
                                         (uint64_t)(ty0 + (int)dy) * g.nz + (uint64_t)(tz0 + (int)dz);
                    unsafeAtomicAdd(b.fluence + idx, v);
                }
            }
        }
        if (tid < 64) {
            if (tid == 0) atomicSub(b.active + job, 1u);
            const int nj = pick_job(b);
            if (tid == 0) {
                if (nj >= 0) atomicAdd(b.active + nj, 1u);
                s_job = nj;
            }
        }
        __syncthreads();
    }
}

__global__ void fill_zero_f32(float* p, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0.f;
}
// zero the per-job energies, cursors, workgroup counts and detection counters
__global__ void zero_jobs(double* energy, unsigned long long* jcur, uint32_t* active, uint32_t* ndet, uint32_t njobs) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < njobs; i += gridDim.x * blockDim.x) {
        energy[3 * i] = 0.0;
        energy[3 * i + 1] = 0.0;
        energy[3 * i + 2] = 0.0;
        jcur[i] = 0ull;
        active[i] = 0u;
        ndet[i] = 0u;
    }
}

}  // namespace dev

namespace {
template <typename T>
T* dalloc_b(size_t n) {
    void* p = nullptr;
    if (n == 0) n = 1;
    HIP_OK(hipMalloc(&p, n * sizeof(T)));
    return (T*)p;
}

constexpr uint64_t kPilotPhotons = 16384;     // counting pilot: photons per job (at most)
constexpr uint64_t kTrialDiv = 32;            // trial launches run 1/32 of every job ...
constexpr uint64_t kTrialMin = 4096;          // ... but at least this many photons
constexpr double kCostBase = 4.0;             // work per photon beside its deposits (launch, scatters)
constexpr size_t kLdsFixed = 32 * 16 + dev::kMaxBatchMedia * 16;
constexpr int kMaxTileCells = 40960;

}  // namespace

BatchSession::BatchSession(const std::vector<Problem>& jobs, const LaunchConfig& lc) : probs_(jobs), lc_(lc) {
    if (jobs.empty()) throw std::runtime_error("a batch needs at least one job");
    if (device_count() <= 0) throw std::runtime_error("no HIP device available: the umcx GPU path needs an MI355X (gfx950)");
    HIP_OK(hipSetDevice(lc.device));
    hipStream_t s;
    HIP_OK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    stream_ = s;
    hipEvent_t e0, e1;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    ev0_ = e0;
    ev1_ = e1;

    // concatenated labels and media
    std::vector<float> media;
    uint64_t lab_total = 0;
    for (const Problem& p : jobs) lab_total += (uint64_t)p.nx * p.ny * p.nz;
    std::vector<uint8_t> labels(lab_total);
    std::vector<float> detpos;
    uint64_t lab_off = 0;
    flu_off_.resize(jobs.size());
    for (size_t j = 0; j < jobs.size(); j++) {
        const Problem& p = jobs[j];
        if (p.mediumnum > (int)dev::kMaxBatchMedia) throw std::runtime_error("batched jobs are limited to 128 media each");
        if (p.nphoton > 0xFFFFFFFFull) throw std::runtime_error("batched jobs are limited to 2^32 photons each");
        const uint64_t nxyz = (uint64_t)p.nx * p.ny * p.nz;
        for (uint64_t i = 0; i < nxyz; i++) {
            uint32_t v = (uint32_t)p.labels[i];
            labels[lab_off + i] = (uint8_t)((v & kMedMask) | ((v & kDetMask) ? 0x80u : 0u));
        }
        media.insert(media.end(), p.media.begin(), p.media.end());
        // detected photons: this job's detectors and row slots
        JobRows jr;
        if (p.issavedet && p.detnum > 0) {
            jr.det_off = (uint32_t)(detpos.size() / 4);
            jr.detnum = (uint32_t)p.detnum;
            detpos.insert(detpos.end(), p.detpos.begin(), p.detpos.end());
            jr.row_off = rows_total_;
            jr.detp_off = detp_floats_;
            jr.maxdet = p.maxdetphotons;
            jr.detlen = (uint32_t)p.detp_columns();
            jr.ppathlen = (uint32_t)p.ppath_columns();
            detp_floats_ += (uint64_t)jr.maxdet * jr.detlen;
            rows_total_ += jr.maxdet;
            pp_max_ = std::max(pp_max_, jr.ppathlen);
            any_det_ = true;
        }
        rows_.push_back(jr);
        lab_off += nxyz;
        flu_off_[j] = flu_total_;
        flu_total_ += p.issavevol ? nxyz * (uint64_t)p.maxgate : 0;
    }
    d_labels_ = dalloc_b<uint8_t>(labels.size());
    HIP_OK(hipMemcpy(d_labels_, labels.data(), labels.size(), hipMemcpyHostToDevice));
    d_media_ = dalloc_b<float>(media.size());
    HIP_OK(hipMemcpy(d_media_, media.data(), media.size() * 4, hipMemcpyHostToDevice));
    d_jobs_ = dalloc_b<dev::JobDesc>(jobs.size());
    d_fluence_ = dalloc_b<float>(flu_total_);
    d_energy_ = dalloc_b<double>(3 * jobs.size());
    d_jcur_ = dalloc_b<unsigned long long>(jobs.size());
    d_active_ = dalloc_b<uint32_t>(jobs.size());
    d_ndet_ = dalloc_b<uint32_t>(jobs.size());
    d_detpos_ = dalloc_b<float>(std::max<size_t>(detpos.size(), 4));
    if (!detpos.empty()) HIP_OK(hipMemcpy(d_detpos_, detpos.data(), detpos.size() * 4, hipMemcpyHostToDevice));
    d_detp_ = dalloc_b<float>(detp_floats_);
    d_detid_ = dalloc_b<uint32_t>(rows_total_);
    d_detw_ = dalloc_b<float>(rows_total_);
    boxes_.assign(jobs.size(), TallyBox{0, 0, 0, 0, 0, 0, 0, 0});
    cost_.assign(jobs.size(), 1.0);

    hipDeviceProp_t prop;
    HIP_OK(hipGetDeviceProperties(&prop, lc.device));
    cus_ = prop.multiProcessorCount;
    first_cap_ = cus_ * 32;
    d_first_ = dalloc_b<int32_t>((size_t)first_cap_);

    tune();

    // final job table and initial assignment; capture zero-fill + persistent launch once
    std::vector<double> work;
    upload_jobs(chosen_.tile_cells ? &boxes_ : nullptr, 0, 0, 0, work);
    upload_first(chosen_.blocks, work);
    dev::BatchArgs ba;
    ba.jobs = (const dev::JobDesc*)d_jobs_;
    ba.njobs = (uint32_t)jobs.size();
    ba.tile_cap = (uint32_t)chosen_.tile_cells;
    ba.first = d_first_;
    ba.labels = (const uint8_t*)d_labels_;
    ba.media = (const float4*)d_media_;
    ba.fluence = d_fluence_;
    ba.energy = d_energy_;
    ba.jcur = d_jcur_;
    ba.active = d_active_;
    ba.count_mode = 0;
    set_rows(&ba);
    const size_t lds = lds_fixed(chosen_.block) + (size_t)chosen_.tile_cells * 4;
    hipGraph_t graph;
    HIP_OK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    if (flu_total_) hipLaunchKernelGGL(dev::fill_zero_f32, dim3(1024), dim3(256), 0, s, d_fluence_, flu_total_);
    hipLaunchKernelGGL(dev::zero_jobs, dim3(4), dim3(256), 0, s, d_energy_, d_jcur_, d_active_, d_ndet_,
                       (uint32_t)jobs.size());
    if (any_det_) hipLaunchKernelGGL(dev::batch_kernel<true>, dim3(chosen_.blocks), dim3(chosen_.block), lds, s, ba);
    else hipLaunchKernelGGL(dev::batch_kernel<false>, dim3(chosen_.blocks), dim3(chosen_.block), lds, s, ba);
    HIP_OK(hipStreamEndCapture(s, &graph));
    graph_ = graph;
    hipGraphExec_t exec;
    HIP_OK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    graph_exec_ = exec;
    zero_state(true);
}

BatchSession::~BatchSession() {
    (void)hipSetDevice(lc_.device);
    if (graph_exec_) (void)hipGraphExecDestroy((hipGraphExec_t)graph_exec_);
    if (graph_) (void)hipGraphDestroy((hipGraph_t)graph_);
    for (void* ptr : {d_labels_, d_media_, d_jobs_, (void*)d_first_, (void*)d_jcur_, (void*)d_active_,
                      (void*)d_fluence_, (void*)d_energy_, (void*)d_ndet_, (void*)d_detpos_, (void*)d_detp_,
                      (void*)d_detid_, (void*)d_detw_})
        if (ptr) (void)hipFree(ptr);
    if (ev0_) (void)hipEventDestroy((hipEvent_t)ev0_);
    if (ev1_) (void)hipEventDestroy((hipEvent_t)ev1_);
    if (stream_) (void)hipStreamDestroy((hipStream_t)stream_);
}

// Job table with photon counts n -> clamp(n / div, min, max) when div > 0 (pilot and trials),
// tally boxes from `boxes` (none when null); work[j] = photons x cost, the scheduling weight.
void BatchSession::upload_jobs(const std::vector<TallyBox>* boxes, uint64_t div, uint64_t mn, uint64_t mx,
                               std::vector<double>& work) {
    std::vector<dev::JobDesc> descs(probs_.size());
    work.assign(probs_.size(), 0.0);
    uint64_t lab_off = 0;
    uint32_t med_off = 0;
    for (size_t j = 0; j < probs_.size(); j++) {
        const Problem& p = probs_[j];
        dev::JobDesc& d = descs[j];
        memset(&d, 0, sizeof d);
        d.srcpos = make_float4(p.srcpos[0], p.srcpos[1], p.srcpos[2], p.srcpos[3]);
        d.srcdir = make_float4(p.srcdir[0], p.srcdir[1], p.srcdir[2], p.srcdir[3]);
        d.srcp1 = make_float4(p.srcparam1[0], p.srcparam1[1], p.srcparam1[2], p.srcparam1[3]);
        d.srcp2 = make_float4(p.srcparam2[0], p.srcparam2[1], p.srcparam2[2], p.srcparam2[3]);
        d.nx = p.nx; d.ny = p.ny; d.nz = p.nz; d.nyz = p.ny * p.nz;
        d.nxyz = (uint64_t)p.nx * p.ny * p.nz;
        d.label_off = lab_off;
        d.flu_off = flu_off_[j];
        uint64_t n = p.nphoton;
        if (div) n = std::min<uint64_t>(n, std::min<uint64_t>(mx, std::max<uint64_t>(mn, n / div)));
        d.nphoton = n;
        for (int i = 0; i < 4; i++) d.seeds[i] = p.seeds[i];
        d.tstart = p.tstart; d.tend = p.tend; d.rtstep = p.rtstep; d.maxgate = p.maxgate;
        d.media_off = med_off;
        d.mediumnum = (uint32_t)p.mediumnum;
        d.srctype = p.srctype; d.isreflect = p.isreflect != 0; d.outputtype = p.outputtype; d.issavevol = p.issavevol;
        if (boxes) {
            const TallyBox& bx = (*boxes)[j];
            d.tx0 = bx[0]; d.ty0 = bx[1]; d.tz0 = bx[2]; d.tg0 = bx[3];
            d.tsx = bx[4]; d.tsy = bx[5]; d.tsz = bx[6]; d.tsg = bx[7];
        }
        d.cost = (float)(cost_[j] + kCostBase);
        const JobRows& jr = rows_[j];
        d.detnum = jr.detnum; d.det_off = jr.det_off; d.maxdet = jr.maxdet; d.row_off = jr.row_off;
        d.detp_off = jr.detp_off; d.detlen = jr.detlen; d.ppathlen = jr.ppathlen; d.savedetflag = p.savedetflag;
        work[j] = (double)n * d.cost;
        lab_off += d.nxyz;
        med_off += (uint32_t)p.mediumnum;
    }
    HIP_OK(hipMemcpy(d_jobs_, descs.data(), descs.size() * sizeof(dev::JobDesc), hipMemcpyHostToDevice));
}

// Initial job of every workgroup: workgroups are shared out in proportion to each job's work
// (largest remainders; every workgroup gets a job), and a job's workgroups are placed on one
// XCD where they fit (workgroup w runs on XCD w mod 8), so they share that XCD's L2.
void BatchSession::upload_first(int blocks, const std::vector<double>& work) {
    if (blocks > first_cap_) throw std::runtime_error("too many workgroups for the batched launch");
    const size_t nj = work.size();
    double total = 0;
    for (double w : work) total += w;
    std::vector<int> cnt(nj, 0);
    std::vector<std::pair<double, int>> frac;
    int given = 0;
    for (size_t j = 0; j < nj; j++) {
        const double share = total > 0 ? blocks * work[j] / total : 0.0;
        cnt[j] = (int)share;
        given += cnt[j];
        if (work[j] > 0) frac.push_back({share - cnt[j], (int)j});
    }
    std::sort(frac.begin(), frac.end(), [](const auto& a, const auto& b) { return a.first > b.first || (a.first == b.first && a.second < b.second); });
    for (size_t i = 0; given < blocks && !frac.empty(); i = (i + 1) % frac.size(), given++) cnt[frac[i].second]++;
    std::vector<int> order;
    order.reserve(blocks);
    for (size_t j = 0; j < nj; j++)
        for (int k = 0; k < cnt[j]; k++) order.push_back((int)j);
    while ((int)order.size() < blocks) order.push_back(0);
    std::vector<int32_t> first(blocks);
    const int xcds = 8;
    if (blocks % xcds == 0) {
        const int per = blocks / xcds;
        for (int i = 0; i < blocks; i++) first[(i % per) * xcds + i / per] = order[i];
    } else {
        for (int i = 0; i < blocks; i++) first[i] = order[i];
    }
    HIP_OK(hipMemcpy(d_first_, first.data(), first.size() * 4, hipMemcpyHostToDevice));
}

void BatchSession::zero_state(bool fluence) {
    hipStream_t s = (hipStream_t)stream_;
    if (fluence && flu_total_) hipLaunchKernelGGL(dev::fill_zero_f32, dim3(1024), dim3(256), 0, s, d_fluence_, flu_total_);
    hipLaunchKernelGGL(dev::zero_jobs, dim3(4), dim3(256), 0, s, d_energy_, d_jcur_, d_active_, d_ndet_,
                       (uint32_t)probs_.size());
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(s));
}

double BatchSession::launch_once(int block, int blocks, int tile_cells, const std::vector<TallyBox>* boxes,
                                 uint64_t div, uint64_t mn, uint64_t mx, int count_mode) {
    hipStream_t s = (hipStream_t)stream_;
    std::vector<double> work;
    upload_jobs(boxes, div, mn, mx, work);
    upload_first(blocks, work);
    zero_state(count_mode != 0);
    dev::BatchArgs ba;
    ba.jobs = (const dev::JobDesc*)d_jobs_;
    ba.njobs = (uint32_t)probs_.size();
    ba.tile_cap = (uint32_t)tile_cells;
    ba.first = d_first_;
    ba.labels = (const uint8_t*)d_labels_;
    ba.media = (const float4*)d_media_;
    ba.fluence = d_fluence_;
    ba.energy = d_energy_;
    ba.jcur = d_jcur_;
    ba.active = d_active_;
    ba.count_mode = count_mode;
    set_rows(&ba);
    // the counting pilot needs no detected photons; trials time the kernel the runs use
    const bool det = any_det_ && !count_mode;
    const size_t lds = (det ? lds_fixed(block) : kLdsFixed) + (size_t)tile_cells * 4;
    HIP_OK(hipEventRecord((hipEvent_t)ev0_, s));
    if (det) hipLaunchKernelGGL(dev::batch_kernel<true>, dim3(blocks), dim3(block), lds, s, ba);
    else hipLaunchKernelGGL(dev::batch_kernel<false>, dim3(blocks), dim3(block), lds, s, ba);
    HIP_OK(hipGetLastError());
    HIP_OK(hipEventRecord((hipEvent_t)ev1_, s));
    HIP_OK(hipStreamSynchronize(s));
    float ms = 0.f;
    HIP_OK(hipEventElapsedTime(&ms, (hipEvent_t)ev0_, (hipEvent_t)ev1_));
    return ms;
}

// Autotuner: counting pilot -> per-job cost and deposit maps; candidate launch shapes (block
// size x workgroups per CU, which fixes each workgroup's LDS share and so its tile size) are
// timed on a sample of every job; the fastest wins. Explicit launch settings narrow the search.
void BatchSession::tune() {
    hipStream_t s = (hipStream_t)stream_;
    (void)s;
    hipDeviceProp_t prop;
    HIP_OK(hipGetDeviceProperties(&prop, lc_.device));
    const size_t lds_cu = prop.maxSharedMemoryPerMultiProcessor ? prop.maxSharedMemoryPerMultiProcessor : 160 * 1024;
    const size_t lds_wg = prop.sharedMemPerBlock ? prop.sharedMemPerBlock : 64 * 1024;
    const void* kfn = any_det_ ? (const void*)dev::batch_kernel<true> : (const void*)dev::batch_kernel<false>;
    auto occupancy = [&](int block, size_t lds) {
        int per_cu = 0;
        HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, block, lds));
        return std::max(per_cu, 0);
    };

    // 1. counting pilot (no tile): deposits per photon and the deposit map of every job
    int pocc = 0;
    HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&pocc, (const void*)dev::batch_kernel<false>, 256, kLdsFixed));
    const int pblock = 256, pblocks = cus_ * std::max(1, pocc);
    launch_once(pblock, std::min(pblocks, first_cap_), 0, nullptr, 1, kPilotPhotons, kPilotPhotons, 1);
    std::unique_ptr<float[]> counts(new float[std::max<uint64_t>(flu_total_, 1)]);   // filled by the copy
    if (flu_total_) HIP_OK(hipMemcpy(counts.get(), d_fluence_, flu_total_ * 4, hipMemcpyDeviceToHost));
    std::vector<CountGrid> grids(probs_.size());
    for (size_t j = 0; j < probs_.size(); j++) {
        const Problem& p = probs_[j];
        const uint64_t len = fluence_len((int)j);
        grids[j] = CountGrid{len ? counts.get() + flu_off_[j] : nullptr, (int)p.nx, (int)p.ny, (int)p.nz, len ? p.maxgate : 0};
    }
    const std::vector<double> totals = grid_totals(grids);
    for (size_t j = 0; j < probs_.size(); j++) {
        const uint64_t n = std::min<uint64_t>(probs_[j].nphoton, kPilotPhotons);
        cost_[j] = n ? totals[j] / (double)n : 0.0;
    }

    // 2. candidates
    std::vector<int> blocks_list = {256, 512, 1024};
    if (lc_.block > 0) blocks_list = {std::max(64, std::min(1024, (lc_.block / 64) * 64))};
    std::vector<BatchTrial> cands;
    for (int block : blocks_list) {
        const size_t fixed = lds_fixed(block);
        if (fixed > lds_wg) continue;
        const int occ = occupancy(block, fixed);
        if (occ < 1) continue;
        for (int per_cu = occ; per_cu >= std::max(1, occ - 1); per_cu--) {
            BatchTrial t;
            t.block = block;
            t.blocks = lc_.blocks > 0 ? lc_.blocks : cus_ * per_cu;
            const int wg_cu = lc_.blocks > 0 ? (lc_.blocks + cus_ - 1) / cus_ : per_cu;
            const size_t share = std::min(lds_wg, lds_cu / (size_t)std::max(1, wg_cu));
            int cells = share > fixed ? (int)std::min<size_t>((share - fixed) / 4, kMaxTileCells) : 0;
            cells = (cells / 256) * 256;
            if (lc_.tally_mode == 0) cells = 0;
            // largest tile that keeps the wanted workgroups per CU resident
            while (cells > 0 && occupancy(block, fixed + (size_t)cells * 4) < std::min(per_cu, occ)) cells -= 256;
            t.tile_cells = cells;
            bool dup = false;
            for (const BatchTrial& c : cands) dup |= c.block == t.block && c.blocks == t.blocks && c.tile_cells == t.tile_cells;
            if (!dup) cands.push_back(t);
            if (lc_.blocks > 0) break;
        }
    }
    if (cands.empty()) throw std::runtime_error("no launch shape fits the batched kernel");

    // 3. tally boxes per distinct tile size
    std::vector<int> sizes;
    for (const BatchTrial& c : cands)
        if (c.tile_cells && std::find(sizes.begin(), sizes.end(), c.tile_cells) == sizes.end()) sizes.push_back(c.tile_cells);
    std::vector<std::vector<TallyBox>> box_sets(sizes.size());
    std::vector<double> shares(sizes.size(), 0.0);
    double all_dep = 0;
    for (size_t j = 0; j < probs_.size(); j++) all_dep += cost_[j] * (double)probs_[j].nphoton;
    for (size_t k = 0; k < sizes.size(); k++) {
        box_sets[k] = choose_boxes(grids, (uint64_t)sizes[k]);
        const std::vector<double> in_box = box_counts(grids, box_sets[k]);
        double in = 0;
        for (size_t j = 0; j < probs_.size(); j++)
            if (totals[j] > 0) in += cost_[j] * (double)probs_[j].nphoton * (in_box[j] / totals[j]);
        shares[k] = all_dep > 0 ? in / all_dep : 0.0;
    }
    auto set_of = [&](int cells) -> const std::vector<TallyBox>* {
        for (size_t k = 0; k < sizes.size(); k++)
            if (sizes[k] == cells) return &box_sets[k];
        return nullptr;
    };
    auto share_of = [&](int cells) {
        for (size_t k = 0; k < sizes.size(); k++)
            if (sizes[k] == cells) return shares[k];
        return 0.0;
    };

    // 4. trials on a sample of every job. Shapes with an LDS tile win over tile-less ones unless
    // tiles are switched off: a tile also sums each workgroup's hot-cell deposits apart before
    // one add to HBM, which keeps the float32 tallies of the hottest voxels accurate (a voxel
    // under a pencil source takes ~1e7 single-precision adds otherwise).
    // A single candidate (an explicit shape) needs no trial launch; its time is reported as 0.
    trials_.clear();
    for (const BatchTrial& c : cands) {
        BatchTrial t = c;
        t.ms = cands.size() > 1 ? launch_once(t.block, t.blocks, t.tile_cells, set_of(t.tile_cells), kTrialDiv,
                                              kTrialMin, ~0ull, 0)
                                : 0.0;
        t.tile_share = share_of(t.tile_cells);
        trials_.push_back(t);
    }
    auto better = [](const BatchTrial& a, const BatchTrial& b) {
        if ((a.tile_cells > 0) != (b.tile_cells > 0)) return a.tile_cells > 0;
        return a.ms < b.ms;
    };
    chosen_ = *std::min_element(trials_.begin(), trials_.end(), better);
    if (chosen_.tile_cells) boxes_ = *set_of(chosen_.tile_cells);
    else boxes_.assign(probs_.size(), TallyBox{0, 0, 0, 0, 0, 0, 0, 0});
}

RunStats BatchSession::run() {
    HIP_OK(hipSetDevice(lc_.device));
    hipStream_t s = (hipStream_t)stream_;
    RunStats st;
    for (const Problem& p : probs_) st.photons += p.nphoton;
    HIP_OK(hipEventRecord((hipEvent_t)ev0_, s));
    HIP_OK(hipGraphLaunch((hipGraphExec_t)graph_exec_, s));
    HIP_OK(hipEventRecord((hipEvent_t)ev1_, s));
    HIP_OK(hipStreamSynchronize(s));
    float ms = 0.f;
    HIP_OK(hipEventElapsedTime(&ms, (hipEvent_t)ev0_, (hipEvent_t)ev1_));
    st.kernel_ms = ms;
    return st;
}

uint64_t BatchSession::fluence_len(int j) const {
    const Problem& p = probs_.at(j);
    return p.issavevol ? (uint64_t)p.nx * p.ny * p.nz * (uint64_t)p.maxgate : 0;
}

void BatchSession::fetch_fluence(int j, float* out) const {
    HIP_OK(hipSetDevice(lc_.device));
    uint64_t n = fluence_len(j);
    if (n) HIP_OK(hipMemcpy(out, d_fluence_ + flu_off_.at(j), n * 4, hipMemcpyDeviceToHost));
}

// Dynamic LDS beside the tally tile: libm tables and media, plus the per-lane partial paths when
// any job saves detected photons.
size_t BatchSession::lds_fixed(int block) const {
    return kLdsFixed + (any_det_ ? (size_t)pp_max_ * (size_t)block * 4 : 0);
}

void BatchSession::set_rows(void* args) const {
    dev::BatchArgs& ba = *(dev::BatchArgs*)args;
    ba.detpos = (const float4*)d_detpos_;
    ba.detp = d_detp_;
    ba.detid = d_detid_;
    ba.detw = d_detw_;
    ba.ndet = d_ndet_;
    ba.pp_max = pp_max_;
}

uint32_t BatchSession::detected(int j) const {
    HIP_OK(hipSetDevice(lc_.device));
    if (j < 0 || j >= njobs()) throw std::runtime_error("job index out of range");
    uint32_t n = 0;
    HIP_OK(hipMemcpy(&n, d_ndet_ + j, 4, hipMemcpyDeviceToHost));
    return n;
}

uint32_t BatchSession::saved(int j) const { return std::min(detected(j), rows_.at(j).maxdet); }

void BatchSession::fetch_detp(int j, float* rows, uint32_t* ids, float* weights) const {
    HIP_OK(hipSetDevice(lc_.device));
    const JobRows& jr = rows_.at(j);
    const uint32_t n = saved(j);
    if (rows && n && jr.detlen)
        HIP_OK(hipMemcpy(rows, d_detp_ + jr.detp_off, (size_t)n * jr.detlen * 4, hipMemcpyDeviceToHost));
    if (ids && n) HIP_OK(hipMemcpy(ids, d_detid_ + jr.row_off, (size_t)n * 4, hipMemcpyDeviceToHost));
    if (weights && n) HIP_OK(hipMemcpy(weights, d_detw_ + jr.row_off, (size_t)n * 4, hipMemcpyDeviceToHost));
}

void BatchSession::energies(int j, double out[3]) const {
    HIP_OK(hipSetDevice(lc_.device));
    if (j < 0 || j >= njobs()) throw std::runtime_error("job index out of range");
    HIP_OK(hipMemcpy(out, d_energy_ + 3 * j, 24, hipMemcpyDeviceToHost));
}

}  // namespace umcx
