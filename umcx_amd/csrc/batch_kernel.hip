// batch_kernel.hip -- persistent gfx950 kernel for many independent simulations (batch.h).
//
// The per-photon transport is the plain kernel's (photon_device.h); what changes is where a lane
// gets its next photon and its job's constants. A lane keeps its current job's geometry, tally
// and label bases and time-gate constants in registers and reloads them only when its next photon
// belongs to another job. Energy sums are kept per lane and flushed to the job's fp64 slots on a
// job change. Media of all jobs are staged in LDS when they fit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
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

// one job's constants (MCX_param :95-99 + source + seeds + buffer offsets)
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
    uint32_t pad[2];
};

struct BatchArgs {
    const JobDesc* jobs;
    const uint2* chunks;        // {job, first photon} per chunk
    uint32_t chunk;             // photon slots per chunk
    uint64_t slots;             // chunks * chunk
    const uint8_t* labels;      // all jobs' labels, bit 7 = detector mask
    const float4* media;        // all jobs' media
    uint32_t total_media;
    int32_t media_in_lds;
    float* fluence;
    double* energy;             // 3 per job
    unsigned long long* cursor;
};

__device__ __forceinline__ void flush_energy(double* e, int job, double& esc, double& tot, double& dep) {
    if (job >= 0 && (tot != 0.0 || esc != 0.0 || dep != 0.0)) {
        unsafeAtomicAdd(e + 3 * job + 0, esc);
        unsafeAtomicAdd(e + 3 * job + 1, tot);
        unsafeAtomicAdd(e + 3 * job + 2, dep);
    }
    esc = tot = dep = 0.0;
}

__global__ void __launch_bounds__(256) batch_kernel(BatchArgs b) {
    extern __shared__ float4 lds4[];
    libm::LogfEntry* slog = (libm::LogfEntry*)lds4;
    uint64_t* sexp = (uint64_t*)(lds4 + 16);
    float4* smedia = lds4 + 32;
    const int tid = threadIdx.x, bd = blockDim.x;
    if (tid < 16) slog[tid] = libm::kLogfTab[tid];
    if (tid < 32) sexp[tid] = libm::kExp2Tab[tid];
    if (b.media_in_lds)
        for (uint32_t i = tid; i < b.total_media; i += bd) smedia[i] = b.media[i];
    __syncthreads();
    const Tabs tabs{slog, sexp};
    const float4* media = b.media_in_lds ? smedia : b.media;

    Photon p;
    Rng r;
    bool alive = false;
    int job = -1;
    Grid g{0, 0, 0, 0};
    uint64_t lab_off = 0, flu_off = 0, nxyz = 0;
    uint32_t med_off = 0;
    float tstart = 0.f, tend = 0.f, rtstep = 0.f;
    int maxgate = 1, isreflect = 0, energy_out = 0, savevol = 1;
    float w0 = 0.f;
    double e_esc = 0.0, e_tot = 0.0, e_dep = 0.0, ph_dep = 0.0;

    while (true) {
        if (!alive) {
            const unsigned long long need = __ballot(1);
            const int leader = __ffsll((long long)need) - 1;
            unsigned long long base = 0;
            if ((int)__lane_id() == leader) base = atomicAdd(b.cursor, (unsigned long long)__popcll(need));
            base = __shfl(base, leader);
            const unsigned long long k = base + (unsigned long long)__popcll(need & (((1ull << __lane_id()) - 1ull)));
            if (k >= b.slots) break;
            const uint2 c = b.chunks[k / b.chunk];
            const JobDesc& J = b.jobs[c.x];
            const uint64_t pid = (uint64_t)c.y + k % b.chunk;
            if (pid >= J.nphoton) continue;                 // padding slot of a job's last chunk
            if ((int)c.x != job) {
                flush_energy(b.energy, job, e_esc, e_tot, e_dep);
                job = (int)c.x;
                g = Grid{J.nx, J.ny, J.nz, J.nyz};
                lab_off = J.label_off; flu_off = J.flu_off; nxyz = J.nxyz; med_off = J.media_off;
                tstart = J.tstart; tend = J.tend; rtstep = J.rtstep; maxgate = J.maxgate;
                isreflect = J.isreflect; energy_out = J.outputtype == kEnergy; savevol = J.issavevol;
            }
            const uint32_t i32 = (uint32_t)pid;
            rng_seed(r, J.seeds[0] ^ i32, J.seeds[1] | i32, J.seeds[2] ^ i32, J.seeds[3] | i32);   // :761
            const Source src{J.srcpos, J.srcdir, J.srcp1, J.srcp2, J.srctype};
            w0 = launch(p, r, src);
            ph_dep = 0.0;
            p.lastvox = vox_index(g, p.ix, p.iy, p.iz);
            bool entered = p.lastvox >= 0 || skip_to_box(p, g);
            if (entered && p.lastvox < 0) entered = false;
            if (!entered) continue;
            const uint32_t lab = b.labels[lab_off + (uint64_t)p.lastvox];
            p.med = (lab & 0x7Fu) | ((lab & 0x80u) << 24);
            const float4 m = media[med_off + (p.med & kMed)];
            p.cmua = m.x;
            p.cmus = m.y;
            p.lx = rng_scatlen(r, tabs);
            alive = true;
        }

        bool ended = false;
        if (p.lx > 0.f) {
            const float4 prop = media[med_off + (p.med & kMed)];
            int nv = step(p, prop.w, g, tabs);
            if (nv != p.lastvox) {
                const uint32_t mcur = p.med & kMed;
                if (savevol) {
                    int tshift = (int)fminf((float)(maxgate - 1), (float)(int)floorf((p.ly - tstart) * rtstep));
                    float dw = p.lw - p.pw;
                    float val = energy_out ? dw : ((p.cmua < kEps) ? (p.lw * p.lz) : dw / p.cmua);
                    ph_dep += (double)dw;
                    if (tshift >= 0) unsafeAtomicAdd(b.fluence + flu_off + (uint64_t)p.lastvox + (uint64_t)tshift * nxyz, val);
                    p.lw = p.pw;
                    p.lz = 0.f;
                }
                if (p.ly > tend) {
                    ended = true;
                } else {
                    uint32_t nm = 0u;
                    if (nv >= 0) {
                        const uint32_t l2 = b.labels[lab_off + (uint64_t)nv];
                        nm = (l2 & 0x7Fu) | ((l2 & 0x80u) << 24);
                    }
                    const float4 nprop = media[med_off + (nm & kMed)];
                    if (isreflect && prop.w != nprop.w) {
                        if (boundary(p, prop.w, nprop.w, r, nv, nm) && (nv < 0 || (nm & kMed) == 0)) ended = true;
                    } else if (nv < 0 || ((nm & kMed) == 0 && mcur != 0)) {
                        ended = true;
                    }
                    if (!ended) {
                        const bool moved = nv != p.lastvox;
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
            if (!(p.lx > 0.f)) scatter(p, r, media[med_off + (p.med & kMed)], tabs);
            continue;
        }
        if (p.med & kMed) {
            e_tot += (double)w0;
            e_esc += (double)p.pw;
            e_dep += ph_dep;
        }
        alive = false;
    }
    flush_energy(b.energy, job, e_esc, e_tot, e_dep);
}

__global__ void fill_zero_f32(float* p, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0.f;
}
__global__ void fill_zero_f64(double* p, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}
__global__ void zero_cursor(unsigned long long* p) {
    if (threadIdx.x == 0 && blockIdx.x == 0) p[0] = 0ull;
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
}  // namespace

BatchSession::BatchSession(const std::vector<Problem>& jobs, const LaunchConfig& lc, uint32_t chunk)
    : probs_(jobs), lc_(lc), chunk_(chunk) {
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

    // job table, concatenated labels and media, round-robin chunk list
    std::vector<dev::JobDesc> descs(jobs.size());
    std::vector<float> media;
    uint64_t lab_total = 0;
    for (const Problem& p : jobs) lab_total += (uint64_t)p.nx * p.ny * p.nz;
    std::vector<uint8_t> labels(lab_total);
    uint64_t lab_off = 0;
    flu_off_.resize(jobs.size());
    for (size_t j = 0; j < jobs.size(); j++) {
        const Problem& p = jobs[j];
        if (p.mediumnum > 128) throw std::runtime_error("batched jobs are limited to 128 media each");
        if (p.issavedet && p.detnum > 0)
            throw std::runtime_error("batched jobs tally fluence only; set DoPartialPath 0 or drop the detectors");
        dev::JobDesc& d = descs[j];
        memset(&d, 0, sizeof d);
        d.srcpos = make_float4(p.srcpos[0], p.srcpos[1], p.srcpos[2], p.srcpos[3]);
        d.srcdir = make_float4(p.srcdir[0], p.srcdir[1], p.srcdir[2], p.srcdir[3]);
        d.srcp1 = make_float4(p.srcparam1[0], p.srcparam1[1], p.srcparam1[2], p.srcparam1[3]);
        d.srcp2 = make_float4(p.srcparam2[0], p.srcparam2[1], p.srcparam2[2], p.srcparam2[3]);
        d.nx = p.nx; d.ny = p.ny; d.nz = p.nz; d.nyz = p.ny * p.nz;
        d.nxyz = (uint64_t)p.nx * p.ny * p.nz;
        d.label_off = lab_off;
        d.flu_off = flu_total_;
        flu_off_[j] = flu_total_;
        d.nphoton = p.nphoton;
        for (int i = 0; i < 4; i++) d.seeds[i] = p.seeds[i];
        d.tstart = p.tstart; d.tend = p.tend; d.rtstep = p.rtstep; d.maxgate = p.maxgate;
        d.media_off = (uint32_t)(media.size() / 4);
        d.mediumnum = (uint32_t)p.mediumnum;
        d.srctype = p.srctype; d.isreflect = p.isreflect != 0; d.outputtype = p.outputtype; d.issavevol = p.issavevol;
        for (uint64_t i = 0; i < d.nxyz; i++) {
            uint32_t v = (uint32_t)p.labels[i];
            labels[lab_off + i] = (uint8_t)((v & kMedMask) | ((v & kDetMask) ? 0x80u : 0u));
        }
        media.insert(media.end(), p.media.begin(), p.media.end());
        lab_off += d.nxyz;
        flu_total_ += p.issavevol ? d.nxyz * (uint64_t)p.maxgate : 0;
    }
    std::vector<uint2> chunks;
    for (uint64_t round = 0;; round++) {
        bool any = false;
        for (size_t j = 0; j < jobs.size(); j++) {
            uint64_t first = round * chunk_;
            if (first < jobs[j].nphoton) {
                if (first > 0xFFFFFFFFull) throw std::runtime_error("batched jobs are limited to 2^32 photons each");
                chunks.push_back(make_uint2((uint32_t)j, (uint32_t)first));
                any = true;
            }
        }
        if (!any) break;
    }
    slots_ = (uint64_t)chunks.size() * chunk_;
    total_media_ = (int)(media.size() / 4);

    d_jobs_ = dalloc_b<dev::JobDesc>(descs.size());
    HIP_OK(hipMemcpy(d_jobs_, descs.data(), descs.size() * sizeof(dev::JobDesc), hipMemcpyHostToDevice));
    d_chunks_ = dalloc_b<uint2>(chunks.size());
    HIP_OK(hipMemcpy(d_chunks_, chunks.data(), chunks.size() * sizeof(uint2), hipMemcpyHostToDevice));
    d_labels_ = dalloc_b<uint8_t>(labels.size());
    HIP_OK(hipMemcpy(d_labels_, labels.data(), labels.size(), hipMemcpyHostToDevice));
    d_media_ = dalloc_b<float>(media.size());
    HIP_OK(hipMemcpy(d_media_, media.data(), media.size() * 4, hipMemcpyHostToDevice));
    d_fluence_ = dalloc_b<float>(flu_total_);
    d_energy_ = dalloc_b<double>(3 * jobs.size());
    d_cursor_ = dalloc_b<unsigned long long>(1);

    hipDeviceProp_t prop;
    HIP_OK(hipGetDeviceProperties(&prop, lc.device));
    block_ = lc.block > 0 ? std::max(64, std::min(256, (lc.block / 64) * 64)) : 256;
    const size_t media_bytes = (size_t)total_media_ * 16;
    media_in_lds_ = media_bytes <= 48 * 1024;
    lds_ = 32 * 16 + (media_in_lds_ ? media_bytes : 0);
    int per_cu = 0;
    HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)dev::batch_kernel, block_, lds_));
    if (per_cu < 1) per_cu = 1;
    blocks_ = lc.blocks > 0 ? lc.blocks : prop.multiProcessorCount * per_cu;

    // capture zero-fill + persistent launch once; every run replays the graph
    dev::BatchArgs ba;
    ba.jobs = (const dev::JobDesc*)d_jobs_;
    ba.chunks = (const uint2*)d_chunks_;
    ba.chunk = chunk_;
    ba.slots = slots_;
    ba.labels = (const uint8_t*)d_labels_;
    ba.media = (const float4*)d_media_;
    ba.total_media = (uint32_t)total_media_;
    ba.media_in_lds = media_in_lds_ ? 1 : 0;
    ba.fluence = d_fluence_;
    ba.energy = d_energy_;
    ba.cursor = d_cursor_;
    hipGraph_t graph;
    HIP_OK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    if (flu_total_) hipLaunchKernelGGL(dev::fill_zero_f32, dim3(1024), dim3(256), 0, s, d_fluence_, flu_total_);
    hipLaunchKernelGGL(dev::fill_zero_f64, dim3(1), dim3(256), 0, s, d_energy_, (uint64_t)(3 * jobs.size()));
    hipLaunchKernelGGL(dev::zero_cursor, dim3(1), dim3(64), 0, s, d_cursor_);
    hipLaunchKernelGGL(dev::batch_kernel, dim3(blocks_), dim3(block_), lds_, s, ba);
    HIP_OK(hipStreamEndCapture(s, &graph));
    graph_ = graph;
    hipGraphExec_t exec;
    HIP_OK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    graph_exec_ = exec;
}

BatchSession::~BatchSession() {
    (void)hipSetDevice(lc_.device);
    if (graph_exec_) (void)hipGraphExecDestroy((hipGraphExec_t)graph_exec_);
    if (graph_) (void)hipGraphDestroy((hipGraph_t)graph_);
    for (void* ptr : {d_labels_, d_media_, d_jobs_, d_chunks_, (void*)d_fluence_, (void*)d_energy_, (void*)d_cursor_})
        if (ptr) (void)hipFree(ptr);
    if (ev0_) (void)hipEventDestroy((hipEvent_t)ev0_);
    if (ev1_) (void)hipEventDestroy((hipEvent_t)ev1_);
    if (stream_) (void)hipStreamDestroy((hipStream_t)stream_);
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

void BatchSession::energies(int j, double out[3]) const {
    HIP_OK(hipSetDevice(lc_.device));
    if (j < 0 || j >= njobs()) throw std::runtime_error("job index out of range");
    HIP_OK(hipMemcpy(out, d_energy_ + 3 * j, 24, hipMemcpyDeviceToHost));
}

}  // namespace umcx
