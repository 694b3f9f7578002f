// photon_kernel.hip -- gfx950 photon-transport kernel (plain form) and its Session runtime.
//
// One lane owns one photon at a time. Waves are persistent: the lanes of a wave whose photons
// ended take new photon slots together -- one LDS atomic on the workgroup's static share, or one
// device atomic on the shared tail of the range -- so lanes stay busy while their neighbours'
// photons live on. Photon i is a pure function of (seed words, i, problem) -- its xorshift128+
// stream is reseeded from i (src/umcx.cpp:761) -- so the order in which lanes take photons does
// not change any photon's trajectory.
//
// Arithmetic follows the reference's FP32 operation order exactly (no contraction: built with
// -ffp-contract=off; division and sqrt correctly rounded; denormals kept), and the
// libm calls use the bit-exact replicas of glibc_libm.h -- logf/sinf/cosf/acosf steer the photon;
// expf (src/umcx.cpp:326) only decays its weight, and is replicated too so that weights and
// fluence deposits are also bitwise equal to the reference's, up to atomic summation order.
//
// Fluence deposits go to an LDS tally box (chosen from a counting pilot), else through a
// per-lane cache of the last four global deposits, else straight to HBM as float atomics, which
// execute at the memory side and are the kernel's binding ceiling.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>

#include "../../include/umcx.h"
#include "glibc_libm.h"
#include "photon_device.h"
#include "runtime.h"

#define HIP_OK(x)                                                                                   \
    do {                                                                                            \
        hipError_t e_ = (x);                                                                        \
        if (e_ != hipSuccess)                                                                       \
            throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_) + " at " + \
                                     __FILE__ + ":" + std::to_string(__LINE__));                   \
    } while (0)

namespace umcx {

constexpr uint64_t kLabelCacheBytes = 96 * 1024;   // largest packed label volume staged in LDS
constexpr uint64_t kMaxTileCells = 40960;          // largest LDS tally tile
constexpr int kRegPaths = 4;                       // partial paths kept in registers up to this many media
// below this many photons per lane of the full launch shape a run ends with its slowest photons,
// and it runs on half the lanes so that fewer waves share each SIMD
constexpr uint64_t kShortRunPhotonsPerLane = 12;
constexpr int kHbmLabelWgPerCu = 4;                // workgroups per CU when labels are read from HBM
constexpr int kDepCache = 4;                       // global deposits cached per lane
// lanes whose photons ended wait until at least this many of their wave's lanes need a new photon
// (or no lane of the wave has one left) before the wave takes new photon slots
constexpr int kRefillMin = 2;
constexpr double kCacheBelowShare = 0.5;           // ... used when the LDS tile takes less of the deposits
constexpr double kCacheMinHits = 0.05;             // ... and when at least this share of the pilot's deposits hit it
// Deposit log: sessions whose global deposits would saturate the memory-side atomics append them
// as {cell, value} records instead, and separate passes sum the records into the grid.
constexpr uint32_t kLogSeg = 4096;                 // records per log segment (32 KiB), one wave's at a time
constexpr uint32_t kLogNone = 0xFFFFFFFFu;         // a wave that holds no segment yet
constexpr uint32_t kLogFull = 0xFFFFFFFEu;         // the log is full: the wave's deposits go to atomics
constexpr uint32_t kTileSegs = 2;                 // log segments the scatter pass sorts together
// Records a log holds at most: positions in the passes are 32-bit, and the reduce pass's
// unrolled reads look up to 8 x 1024 records past a position. Each chunk of a run ends with its
// slowest photons alone on the GPU (about 15 ms for the 200^3 skin model), so larger logs mean
// fewer chunks: 2^32 records take 64 GiB with the grouped copy, a quarter of the MI355X's memory.
constexpr uint64_t kLogMaxRecords = (1ull << 32) - (1ull << 20);
static_assert(kLogMaxRecords + 8 * 1024 < (1ull << 32), "log positions stay 32-bit");
constexpr int kBucketBits = 15;                    // fluence cells per bucket of the reduce pass: 2^15 (128 KiB of LDS)
constexpr uint32_t kMaxBuckets = 4096;             // buckets of the log passes' LDS counters (grids below 2^27 cells)
constexpr double kLogMinRecords = 200.0;           // global deposits per photon from which the log is used
constexpr double kVoteBelowScatters = 0.3;         // pilot scatters per crossing below which scatter rounds wait ...
constexpr int kVoteLanes = 16;                     // ... until this many lanes of a wave need one
constexpr int kLogGridPerCu = 2;                   // workgroups per CU of the histogram and scatter passes
constexpr uint64_t kReduceShare = 1u << 16;         // records per workgroup of the reduce pass, at least
constexpr int kReduceGridPerCu = 8;                 // its workgroups per CU, at most
// the photon kernel's static LDS (slot counter, each wave's log segment and fill), kept out of the
// dynamic LDS a launch may take
constexpr size_t kStaticLds = 256;

namespace dev {

struct KernelArgs {
    uint32_t nx, ny, nz, nyz;
    uint64_t nxyz;
    int32_t maxgate, outputtype, savedetflag, mediumnum, detnum, srctype, issavevol, ppathlen, detlen;
    float tstart, tend, rtstep, pad0;
    float4 srcpos, srcdir, srcp1, srcp2;
    uint32_t seeds[4];
    uint64_t first, count;
    const void* labels;
    const float4* media;
    const float4* detpos;
    float* fluence;
    float* detp;
    uint32_t* detid;
    float* detw;
    uint32_t* ndet;
    uint32_t maxdet;
    double* energy;
    unsigned long long* cursor;
    TraceRec* trace;
    unsigned long long* dep_stats;   // pilot only: {global deposits, deposit-cache misses, scatters}
    // LDS tally tile: box [tx0,tx0+tsx) x [ty0,..) x [tz0,..) over gates [tg0,tg0+tsg)
    int32_t tx0, ty0, tz0, tg0;
    uint32_t tsx, tsy, tsz, tsg, tcells, tile_n;
    int32_t count_mode;      // 1: tally the number of deposits instead of their values
    int32_t dep_cache;       // 1: global deposits go through the per-lane write-back cache
    int32_t vote;            // 0: scatter right after the step that ends a free path; k > 0: a wave
                             // scatters only once k of its lanes (or all of them) wait to scatter
    // per-voxel {mua, mus} in per-voxel-length units overriding the labels' (nullptr: labels only)
    const float2* vprops;
    // replay (Jacobian): photon ids, output slice and coefficient per record; gradients out
    const uint32_t* rids;
    const int32_t* rslice;
    const float* rcoef;
    float* gmua;
    float* gmus;
    uint64_t gstride;
    float unit;              // mm per voxel
    float* rw;               // final weight of each replayed photon (may be null)
    // media ids packed plabel_bits (1, 2 or 4) per voxel, staged in LDS (LAB 1 kernels); code c
    // stands for medium plabel_id0 + c * plabel_step
    const uint32_t* plabels;
    uint32_t plabel_words, plabel_lg, plabel_bits, plabel_id0, plabel_step;
    // photon slots: workgroup w first takes [w*per_wg, (w+1)*per_wg) through an LDS counter,
    // then slots from blocks*per_wg on through the device cursor
    uint64_t per_wg;
    // several sources in one launch, {pos xyz, dir xyz} per source: in replay the source of each
    // record is rsrc[k]; otherwise photon slot p is photon p mod msrc_n of source p / msrc_n
    const int32_t* rsrc;
    const float* rsrcs;
    uint64_t msrc_n;
    // deposit log: global deposits appended as {cell, value bits} to kLogSeg-record segments, a
    // wave filling one segment at a time (null: global deposits are float atomics)
    uint2* dlog;
    uint32_t* dlog_fill;              // records written to each segment
    unsigned long long* dlog_next;    // next free segment
    uint32_t dlog_segs;               // segments in the log
};

// The kernel's arguments read again from the kernarg segment where they are used, through a
// pointer the compiler cannot see is invariant: scalar loads at the point of use instead of
// scalar registers held (and spilled to VGPR lanes) across the whole photon loop.
typedef const __attribute__((address_space(4))) KernelArgs* ArgsRef;
__device__ __forceinline__ ArgsRef reread_args() {
    uint64_t v = (uint64_t)__builtin_amdgcn_kernarg_segment_ptr();
    __asm__ volatile("" : "+s"(v));
    return (ArgsRef)v;
}
__device__ __forceinline__ float4 ld4(const __attribute__((address_space(4))) float4* q) {
    const __attribute__((address_space(4))) float* f = (const __attribute__((address_space(4))) float*)q;
    return make_float4(f[0], f[1], f[2], f[3]);
}

// Append one global deposit per active lane to the deposit log. The lanes of a wave that deposit
// together take consecutive records of the wave's current segment (one LDS read and write by
// the first of them); a wave whose segment cannot take them records how far it filled it and
// takes the next free segment from the device counter. Once the log is full the wave's deposits
// go to the grid as atomics, so a run never loses one. s_seg/s_fill hold each wave's segment
// and fill count in LDS, where every lane of the wave sees the latest values (volatile: the lane
// that reads them may not be the one that wrote them last). The pointers carry the LDS address
// space: volatile accesses through generic pointers compile to flat loads and stores with
// system-coherence bits, each followed by a wait for every vector memory access in flight.
typedef __attribute__((address_space(3))) volatile uint32_t LdsVolatileU32;
__device__ __forceinline__ void log_deposit(LdsVolatileU32* s_seg, LdsVolatileU32* s_fill, uint32_t cell,
                                            float val) {
    const ArgsRef a = reread_args();
    const unsigned long long act = __ballot(1);
    const int leader = __ffsll((long long)act) - 1;
    const uint32_t n = (uint32_t)__popcll(act), rank = (uint32_t)__popcll(act & lanemask_lt());
    const int w = threadIdx.x >> 6;
    uint32_t seg = 0, base = 0;
    if ((int)__lane_id() == leader) {
        seg = s_seg[w];
        base = s_fill[w];
        if (seg != kLogFull && base + n > kLogSeg) {
            if (seg != kLogNone) a->dlog_fill[seg] = base;
            const unsigned long long next = atomicAdd(a->dlog_next, 1ull);
            seg = next < a->dlog_segs ? (uint32_t)next : kLogFull;
            base = 0;
        }
        s_seg[w] = seg;
        s_fill[w] = base + n;
    }
    seg = (uint32_t)__builtin_amdgcn_readlane((int)seg, leader);
    base = (uint32_t)__builtin_amdgcn_readlane((int)base, leader);
    if (seg != kLogFull) a->dlog[(uint64_t)seg * kLogSeg + base + rank] = make_uint2(cell, __float_as_uint(val));
    else unsafeAtomicAdd(a->fluence + cell, val);
}

// Label modes (LAB): 0 labels read from the label volume in HBM; 1 media ids packed in LDS.
constexpr int kLabHbm = 0, kLabLds = 1;
template <int LAB, typename LabelT>
struct LabelPick {
    using T = GlobalLabels<LabelT>;
    static __device__ __forceinline__ T make(const KernelArgs& a, const uint32_t*) { return T{(const LabelT*)a.labels}; }
};
template <typename LabelT>
struct LabelPick<kLabLds, LabelT> {
    using T = LdsLabels;
    static __device__ __forceinline__ T make(const KernelArgs& a, const uint32_t* s) {
        return T{s, a.plabel_lg, a.plabel_bits, a.plabel_id0, a.plabel_step};
    }
};

__device__ __forceinline__ int tile_local(const KernelArgs& a, int ix, int iy, int iz) {
    const unsigned dx = (unsigned)(ix - a.tx0), dy = (unsigned)(iy - a.ty0), dz = (unsigned)(iz - a.tz0);
    return (dx < a.tsx && dy < a.tsy && dz < a.tsz) ? (int)((dx * a.tsy + dy) * a.tsz + dz) : -1;
}

// coefficients of the photon's current voxel: the label's medium, or the per-voxel override of
// mua and mus (the refractive index stays the label's)
__device__ __forceinline__ void set_props(Photon& p, const KernelArgs& a, const float4 m) {
    p.cn = m.w;
    if (a.vprops) {
        const float2 v = a.vprops[p.lastvox];
        p.cmua = v.x;
        p.cmus = v.y;
    } else {
        p.cmua = m.x;
        p.cmus = m.y;
    }
}

// REPLAY: re-trace the photons listed in a.rids (photon i is a pure function of i) and add
// each voxel segment's perturbation-MC terms to the gradient slices instead of tallying.
// DLOG: global deposits go to the deposit log (a.dlog) instead of float atomics.
template <bool REFLECT, bool SAVEDET, typename LabelT, bool REPLAY, int LAB, bool DLOG = false>
__global__ void __launch_bounds__(1024) photon_kernel(KernelArgs a) {
    // LDS: libm tables | media | packed labels (LAB 1) | tally tile | per-lane partial paths
    extern __shared__ float4 lds4[];
    __shared__ uint32_t s_taken;                                 // static photon slots handed out
    __shared__ uint32_t s_lseg[16], s_lfill[16];                 // deposit log: each wave's segment and fill
    static_assert(sizeof(uint32_t) * 33 <= kStaticLds, "static LDS exceeds its reserve");
    libm::LogfEntry* slog = (libm::LogfEntry*)lds4;            // 16 x 16 B
    uint64_t* sexp = (uint64_t*)(lds4 + 16);                     // 32 x 8 B
    float4* smedia = lds4 + 32;
    uint32_t* slab = (uint32_t*)(smedia + a.mediumnum);
    float* stile = (float*)(slab + (LAB == kLabLds ? a.plabel_words : 0u));
    float* sppath = stile + a.tile_n;
    const int tid = threadIdx.x, bd = blockDim.x;
    if (tid < 16) slog[tid] = libm::kLogfTab[tid];
    if (tid < 32) sexp[tid] = libm::kExp2Tab[tid];
    if (tid == 0) s_taken = 0u;
    if (tid < 16) {
        s_lseg[tid] = kLogNone;
        s_lfill[tid] = kLogSeg;
    }
    const Tabs tabs{slog, sexp};
    const Grid grid{a.nx, a.ny, a.nz, a.nyz};
    for (int i = tid; i < a.mediumnum; i += bd) smedia[i] = a.media[i];
    if (LAB == kLabLds)
        for (uint32_t i = tid; i < a.plabel_words; i += bd) slab[i] = a.plabels[i];
    for (int i = tid; i < (int)a.tile_n; i += bd) stile[i] = 0.f;
    // partial path lengths per medium: in registers for up to kRegPaths media, else in LDS
    const bool regpp = a.ppathlen <= kRegPaths;
    if (SAVEDET && !regpp)
        for (int i = tid; i < a.ppathlen * bd; i += bd) sppath[i] = 0.f;
    __syncthreads();
    const auto labs = LabelPick<LAB, LabelT>::make(a, slab);
    // detector flag of the voxel a photon ends in (LDS labels carry media ids only)
    auto det_flag = [&](const Photon& q) -> uint32_t {
        if (LAB != kLabHbm) return load_label<LabelT>(a.labels, q.lastvox) & kDet;
        return q.med & kDet;
    };

    Photon p;
    Rng r;
    bool alive = false;
    uint64_t pid = 0;
    float w0 = 0.f;
    double e_esc = 0.0, e_tot = 0.0, e_dep = 0.0, ph_dep = 0.0;
    float rcoef = 0.f, seg_ns = 0.f;
    float pp0 = 0.f, pp1 = 0.f, pp2 = 0.f, pp3 = 0.f;
    uint64_t rbase = 0, rk = 0;
    const bool savevol = !REPLAY && a.issavevol != 0;
    uint32_t dkey[kDepCache];
    float dsum[kDepCache];
#pragma unroll
    for (int j = 0; j < kDepCache; j++) {
        dkey[j] = kNoKey;
        dsum[j] = 0.f;
    }
    uint32_t nglob = 0, nmiss = 0;   // deposit-cache statistics (reported by the pilot)
    uint32_t nscat = 0;              // scatters of the lane's ended photons (reported by the pilot)
    const bool energy_out = a.outputtype == kEnergy;
    // a global deposit: to the deposit log when the session keeps one, else a float atomic
    auto emit = [&](uint32_t cell, float v) {
        if (DLOG) log_deposit((LdsVolatileU32*)s_lseg, (LdsVolatileU32*)s_lfill, cell, v);
        else unsafeAtomicAdd(a.fluence + cell, v);
    };

    while (true) {
        // ---- take a new photon when this lane's has ended (one atomic per wave) ----
        // The workgroup's static slots come through an LDS counter, so a refill does not wait on
        // a device atomic (whose wait would also cover every fluence atomic still in flight).
        // A refill is a long stretch of scalar and LDS round trips for the few lanes that take
        // part, so a lane whose photon ended waits until kRefillMin lanes of its wave need one
        // (or none of them has a photon left); no photon's trajectory depends on when it starts.
        const unsigned long long dead = __ballot(!alive);
        if (!alive && (__popcll(dead) >= kRefillMin || dead == __ballot(1))) {
            const ArgsRef ca = reread_args();
            Source src{ld4(&ca->srcpos), ld4(&ca->srcdir), ld4(&ca->srcp1), ld4(&ca->srcp2), ca->srctype};
            const unsigned long long need = __ballot(1);
            const int leader = __ffsll((long long)need) - 1;
            const uint32_t n = (uint32_t)__popcll(need), rank = (uint32_t)__popcll(need & lanemask_lt());
            uint32_t sb = 0, nstat = 0;
            unsigned long long gb = 0;
            if ((int)__lane_id() == leader) {
                sb = atomicAdd(&s_taken, n);
                nstat = sb < ca->per_wg ? (uint32_t)min((uint64_t)n, ca->per_wg - sb) : 0u;
                if (nstat < n) gb = atomicAdd(ca->cursor, (unsigned long long)(n - nstat));
            }
            sb = __shfl(sb, leader);
            nstat = __shfl(nstat, leader);
            gb = __shfl(gb, leader);
            const unsigned long long k = rank < nstat ? (uint64_t)blockIdx.x * ca->per_wg + sb + rank
                                                      : (uint64_t)gridDim.x * ca->per_wg + gb + (rank - nstat);
            if (k >= ca->count) break;
            pid = REPLAY ? (uint64_t)ca->rids[k] : ca->first + k;
            if (REPLAY) {
                rk = k;
                rcoef = a.rcoef[k];
                rbase = (uint64_t)a.rslice[k] * a.gstride;
                seg_ns = 0.f;
                if (a.rsrc) {
                    const float* q = a.rsrcs + 6 * (uint64_t)a.rsrc[k];
                    src.srcpos.x = q[0]; src.srcpos.y = q[1]; src.srcpos.z = q[2];
                    src.srcdir.x = q[3]; src.srcdir.y = q[4]; src.srcdir.z = q[5];
                }
            }
            uint64_t ip = pid;                       // photon index within its source
            if (!REPLAY && ca->msrc_n) {
                const uint64_t si = pid / ca->msrc_n;
                ip = pid - si * ca->msrc_n;
                const float* q = ca->rsrcs + 6 * si;
                src.srcpos.x = q[0]; src.srcpos.y = q[1]; src.srcpos.z = q[2];
                src.srcdir.x = q[3]; src.srcdir.y = q[4]; src.srcdir.z = q[5];
            }
            const uint32_t i32 = (uint32_t)ip;
            rng_seed(r, ca->seeds[0] ^ i32, ca->seeds[1] | i32, ca->seeds[2] ^ i32, ca->seeds[3] | i32);   // :761
            w0 = launch(p, r, src);
            ph_dep = 0.0;
            if (SAVEDET) {
                pp0 = pp1 = pp2 = pp3 = 0.f;
                if (!regpp)
                    for (int m = 0; m < a.ppathlen; m++) sppath[m * bd + tid] = 0.f;
            }
            p.lastvox = vox_index(grid, p.ix, p.iy, p.iz);                                           // :262
            bool entered = p.lastvox >= 0 || skip_to_box(p, grid);                                   // :264
            if (entered && p.lastvox < 0) entered = false;   // entry exactly on a far face
            if (!entered) {
                p.med = 0;
                alive = false;
                if (ca->trace) {
                    TraceRec& t = ca->trace[k];
                    t.pos[0] = p.px; t.pos[1] = p.py; t.pos[2] = p.pz; t.pos[3] = p.pw;
                    t.vec[0] = p.vx; t.vec[1] = p.vy; t.vec[2] = p.vz; t.vec[3] = p.vw;
                    t.len[0] = p.lx; t.len[1] = p.ly; t.len[2] = p.lz; t.len[3] = p.lw;
                    t.mediaid = 0; t.lastvoxelidx = p.lastvox; t.w0 = w0;
                }
                continue;
            }
            p.tvox = tile_local(a, p.ix, p.iy, p.iz);
            p.med = labs.decode(labs.raw(p.lastvox));                                                // :269
            set_props(p, a, smedia[p.med & kMed]);
            p.lx = rng_scatlen(r, tabs);                                                             // :270
            alive = true;
        }
        if (!alive) continue;          // waiting for its wave's next refill

        bool ended = false;
        // vote > 0: a scatter round runs only when enough lanes wait to scatter; in between, the
        // waiting lanes idle and the others keep stepping
        bool scatter_round = true;
        if (a.vote > 0) {
            const unsigned long long act = __ballot(1), sc = __ballot(!(p.lx > 0.f));
            scatter_round = __popcll(sc) >= (unsigned)a.vote || sc == act;
        }
        // ---- one step of sprint (:282-311) ----
        if (p.lx > 0.f && (a.vote == 0 || !scatter_round)) {
            uint32_t nlab;
            int nv = step(p, p.cn, grid, tabs, labs, nlab);
            if (nv != p.lastvox) {
                const uint32_t m = p.med & kMed;
                if (REPLAY) {
                    // d w / d mua_v = -w L_v;  d log p / d mus_v = n_v / mus_v - L_v   (mm units)
                    const float L = p.lz * a.unit;
                    const float dmus = (seg_ns > 0.f ? seg_ns * a.unit / p.cmus : 0.f) - L;
                    unsafeAtomicAdd(a.gmua + rbase + (uint64_t)p.lastvox, -rcoef * L);
                    unsafeAtomicAdd(a.gmus + rbase + (uint64_t)p.lastvox, rcoef * dmus);
                    p.lz = 0.f;
                    seg_ns = 0.f;
                }
                if (SAVEDET && a.ppathlen && m > 0) {
                    if (regpp) {
                        pp0 = m == 1 ? pp0 + p.lz : pp0;
                        pp1 = m == 2 ? pp1 + p.lz : pp1;
                        pp2 = m == 3 ? pp2 + p.lz : pp2;
                        pp3 = m == 4 ? pp3 + p.lz : pp3;
                    } else {
                        sppath[(m - 1) * bd + tid] += p.lz;
                    }
                }
                if (savevol) {
                    int tshift = (int)fminf((float)(a.maxgate - 1), (float)(int)floorf((p.ly - a.tstart) * a.rtstep));
                    float dw = p.lw - p.pw;
                    float val = energy_out ? dw : ((p.cmua < kEps) ? (p.lw * p.lz) : dw / p.cmua);
                    ph_dep += (double)dw;
                    if (a.count_mode) val = 1.f;
                    if (tshift >= 0) {
                        const unsigned g = (unsigned)(tshift - a.tg0);
                        if (p.tvox >= 0 && g < a.tsg)
                            __hip_atomic_fetch_add(stile + g * a.tcells + (unsigned)p.tvox, val, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_WORKGROUP);
                        else if (a.dep_cache) {
                            nglob++;
                            nmiss += cached_deposit(emit, dkey, dsum, (uint32_t)p.lastvox + (uint32_t)tshift * (uint32_t)a.nxyz, val);
                        } else {
                            const uint64_t cell = (uint64_t)p.lastvox + (uint64_t)tshift * a.nxyz;
                            if (DLOG) log_deposit((LdsVolatileU32*)s_lseg, (LdsVolatileU32*)s_lfill, (uint32_t)cell, val);   // cells < 2^32
                            else unsafeAtomicAdd(a.fluence + cell, val);
                        }
                    }
                    p.lw = p.pw;
                    p.lz = 0.f;
                }
                if (p.ly > a.tend) {
                    ended = true;
                } else {
                    uint32_t nm = (nv >= 0) ? labs.decode(nlab) : 0u;
                    const float4 nprop = smedia[nm & kMed];
                    if (REFLECT && p.cn != nprop.w) {
                        if (boundary(p, p.cn, nprop.w, r, nv, nm) && (nv < 0 || (nm & kMed) == 0)) ended = true;
                    } else if (nv < 0 || ((nm & kMed) == 0 && m != 0)) {
                        ended = true;
                    }
                    if (!ended) {
                        const bool moved = nv != p.lastvox;
                        if (moved) p.tvox = tile_local(a, p.ix, p.iy, p.iz);
                        p.lastvox = nv;
                        p.med = nm;
                        if (moved) set_props(p, a, nprop);
                    }
                }
            }
        }
        if (!ended) {
            // a free path used up scatters (:272-274); a zero-length new path scatters again on
            // the next pass, before any step, as the reference's loop does
            if (scatter_round && !(p.lx > 0.f)) {
                scatter(p, r, smedia[p.med & kMed], tabs);
                if (REPLAY) seg_ns += 1.f;
            }
            continue;
        }

        if (REPLAY) {
            if (a.rw) a.rw[rk] = p.pw;
            alive = false;
            continue;
        }
        // ---- photon ends: detection (:276-278) and energy tallies (:765-770) ----
        if (SAVEDET && p.ly <= a.tend && det_flag(p)) {
            const ArgsRef cd = reread_args();
            for (int d = 0; d < cd->detnum; d++) {
                const float4 dp = cd->detpos[d];
                float dx = dp.x - p.px, dy = dp.y - p.py, dz = dp.z - p.pz;
                if (dx * dx + dy * dy + dz * dz < dp.w * dp.w) {
                    uint32_t slot = atomicAdd(cd->ndet, 1u);
                    if (slot < cd->maxdet) {
                        float* row = cd->detp + (uint64_t)slot * cd->detlen;
                        int c = 0;
                        if (cd->savedetflag & kDetID) row[c++] = (float)(d + 1);
                        if (regpp) {
                            const float pp[kRegPaths] = {pp0, pp1, pp2, pp3};
                            for (int m = 0; m < a.ppathlen; m++) row[c++] = pp[m];
                        } else {
                            for (int m = 0; m < a.ppathlen; m++) row[c++] = sppath[m * bd + tid];
                        }
                        if (cd->savedetflag & kExitPos) { row[c++] = p.px; row[c++] = p.py; row[c++] = p.pz; }
                        if (cd->savedetflag & kExitDir) { row[c++] = p.vx; row[c++] = p.vy; row[c++] = p.vz; }
                        cd->detid[slot] = (uint32_t)pid;
                        cd->detw[slot] = p.pw;
                    }
                }
            }
        }
        // the pilot counts scatters (re-read from the kernarg segment: an argument kept live across
        // the loop here cost the headline workload 0.8 %)
        if (reread_args()->dep_stats) nscat += (uint32_t)p.vw;
        if (p.med & kMed) {
            e_tot += (double)w0;
            e_esc += (double)p.pw;
            e_dep += ph_dep;
        }
        if (reread_args()->trace) {
            TraceRec& t = reread_args()->trace[pid - reread_args()->first];
            t.pos[0] = p.px; t.pos[1] = p.py; t.pos[2] = p.pz; t.pos[3] = p.pw;
            t.vec[0] = p.vx; t.vec[1] = p.vy; t.vec[2] = p.vz; t.vec[3] = p.vw;
            t.len[0] = p.lx; t.len[1] = p.ly; t.len[2] = p.lz; t.len[3] = p.lw;
            t.mediaid = (int32_t)((p.med & kMed) | det_flag(p)); t.lastvoxelidx = p.lastvox; t.w0 = w0;
        }
        alive = false;
    }

    if (savevol) flush_deposits(emit, dkey, dsum);
    // the wave's last segment: how far it was filled
    if (DLOG && __lane_id() == 0) {
        const uint32_t seg = ((LdsVolatileU32*)s_lseg)[tid >> 6];
        if (seg != kLogNone && seg != kLogFull) reread_args()->dlog_fill[seg] = ((LdsVolatileU32*)s_lfill)[tid >> 6];
    }
    if (a.dep_stats) {
        const double g = wave_sum((double)nglob), m = wave_sum((double)nmiss), k = wave_sum((double)nscat);
        if (__lane_id() == 0) {
            atomicAdd(a.dep_stats, (unsigned long long)g);
            atomicAdd(a.dep_stats + 1, (unsigned long long)m);
            atomicAdd(a.dep_stats + 2, (unsigned long long)k);
        }
    }

    // ---- flush the LDS tally tile ----
    if (a.tile_n) {
        __syncthreads();
        const uint32_t plane = a.tsy * a.tsz;
        for (uint32_t i = tid; i < a.tile_n; i += bd) {
            const float v = stile[i];
            if (v != 0.f) {
                const uint32_t g = i / a.tcells, c = i - g * a.tcells;
                const uint32_t dx = c / plane, rem = c - dx * plane, dy = rem / a.tsz, dz = rem - dy * a.tsz;
                const uint64_t idx = (uint64_t)(a.tg0 + g) * a.nxyz + (uint64_t)(a.tx0 + dx) * a.nyz +
                                     (uint64_t)(a.ty0 + dy) * a.nz + (uint64_t)(a.tz0 + dz);
                unsafeAtomicAdd(a.fluence + idx, v);
            }
        }
    }

    // ---- energies: wave sums, then one fp64 atomic per wave ----
    e_esc = wave_sum(e_esc);
    e_tot = wave_sum(e_tot);
    e_dep = wave_sum(e_dep);
    if (__lane_id() == 0) {
        unsafeAtomicAdd(a.energy + 0, e_esc);
        unsafeAtomicAdd(a.energy + 1, e_tot);
        unsafeAtomicAdd(a.energy + 2, e_dep);
    }
}

__global__ void zero_u64(unsigned long long* p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// Pilot count grids stay on the device: the largest count (counts are >= 0, so their bit patterns
// order like their values), the first cell holding it and the total.
__global__ void count_max_kernel(const float* c, uint64_t n, unsigned* maxbits) {
    unsigned m = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        m = max(m, __float_as_uint(c[i]));
    if (m) atomicMax(maxbits, m);
}

__global__ void count_argmax_sum_kernel(const float* c, uint64_t n, const unsigned* maxbits,
                                        unsigned long long* first, double* total) {
    const unsigned m = *maxbits;
    unsigned long long f = ~0ull;
    double t = 0.0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const float v = c[i];
        t += (double)v;
        if (__float_as_uint(v) == m && i < f) f = i;
    }
    if (f != ~0ull) atomicMin(first, f);
    if (t != 0.0) unsafeAtomicAdd(total, t);
}

// Sum of one value per thread over a workgroup of up to 1024 lanes; every thread gets the sum.
__device__ double workgroup_sum(double v, double* red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double t = 0.0;
    for (int k = 0; k < nw; k++) t += red[k];
    return t;
}

// Counts in the box [l, h) of a [t][x][y][z] count grid, one workgroup.
__device__ double box_sum_wg(const float* cnt, uint64_t nxyz, uint64_t nyz, int nz, const int* l, const int* h,
                             double* red) {
    const uint64_t L3 = h[3] - l[3], L2 = h[2] - l[2], L1 = h[1] - l[1], L0 = h[0] - l[0];
    const uint64_t n = L0 * L1 * L2 * L3;
    double s = 0.0;
    for (uint64_t i = threadIdx.x; i < n; i += blockDim.x) {
        const uint64_t z = i % L3, y = i / L3 % L2, x = i / (L3 * L2) % L1, t = i / (L3 * L2 * L1);
        s += (double)cnt[(l[0] + t) * nxyz + (l[1] + x) * nyz + (l[2] + y) * (uint64_t)nz + (l[3] + z)];
    }
    return workgroup_sum(s, red);
}

// The greedy growth of choose_box (tally_box.cpp) over a count grid left in device memory: start
// at the densest cell, then add the face slab with the most counts per added cell (candidates in
// the order t, x, y, z, low side first; ties keep the first) while the box holds at most `budget`
// cells. Counts are whole numbers, so every sum is exact and each choice is the host's.
// out_box: {x0, y0, z0, g0, sx, sy, sz, sg} (zeros for an empty grid); out_in: counts inside.
__global__ void __launch_bounds__(1024) grow_box_kernel(const float* cnt, int nt, int nx, int ny, int nz, uint64_t budget,
                                                        const unsigned long long* start, int* out_box, double* out_in) {
    __shared__ int lo[4], hi[4], go;
    __shared__ double red[16];
    const uint64_t nyz = (uint64_t)ny * nz, nxyz = nyz * nx;
    const int ext[4] = {nt, nx, ny, nz};
    if (threadIdx.x == 0) {
        const unsigned long long b = *start;
        go = (b != ~0ull && budget > 0 && cnt[b] > 0.f) ? 1 : 0;
        if (go) {
            lo[0] = (int)(b / nxyz);
            lo[1] = (int)(b % nxyz / nyz);
            lo[2] = (int)(b % nyz / nz);
            lo[3] = (int)(b % nz);
            for (int d = 0; d < 4; d++) hi[d] = lo[d] + 1;
        }
    }
    __syncthreads();
    if (!go) {
        if (threadIdx.x < 8) out_box[threadIdx.x] = 0;
        if (threadIdx.x == 0) *out_in = 0.0;
        return;
    }
    while (true) {
        int l0[4], h0[4];
        for (int k = 0; k < 4; k++) { l0[k] = lo[k]; h0[k] = hi[k]; }
        uint64_t cur = 1;
        for (int k = 0; k < 4; k++) cur *= (uint64_t)(h0[k] - l0[k]);
        int bd = -1, bside = 0;
        double bscore = 0.0;
        for (int c = 0; c < 8; c++) {
            const int d = c >> 1, side = c & 1;
            const int idx = side ? h0[d] : l0[d] - 1;
            if (idx < 0 || idx >= ext[d]) continue;
            const uint64_t nc = cur / (uint64_t)(h0[d] - l0[d]) * (uint64_t)(h0[d] - l0[d] + 1);
            if (nc > budget) continue;
            int l[4], h[4];
            for (int k = 0; k < 4; k++) { l[k] = l0[k]; h[k] = h0[k]; }
            l[d] = idx;
            h[d] = idx + 1;
            const double add = box_sum_wg(cnt, nxyz, nyz, nz, l, h, red);
            const double score = add / (double)(nc - cur);
            if (add > 0 && score > bscore) {
                bscore = score;
                bd = d;
                bside = side;
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            if (bd < 0) go = 0;
            else if (bside) hi[bd]++;
            else lo[bd]--;
        }
        __syncthreads();
        if (!go) break;
    }
    int l[4], h[4];
    for (int k = 0; k < 4; k++) { l[k] = lo[k]; h[k] = hi[k]; }
    const double in = box_sum_wg(cnt, nxyz, nyz, nz, l, h, red);
    if (threadIdx.x == 0) {
        const int b[8] = {l[1], l[2], l[3], l[0], h[1] - l[1], h[2] - l[2], h[3] - l[3], h[0] - l[0]};
        for (int k = 0; k < 8; k++) out_box[k] = b[k];
        *out_in = in;
    }
}

// per-voxel {mua, mus} in 1/mm -> per-voxel-length units, as the label media are scaled (:804)
__global__ void vprops_kernel(const float* mua, const float* mus, float unit, float2* out, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = make_float2(mua[i] * unit, mus[i] * unit);
}

// Self-check of the short correctly rounded sequences against the compiler's full ones, over a
// whole input range, counting the inputs where they differ:
//   0: sqrt_ns(x) vs sqrtf_cr(x) for every float in [2^-96, 2^128)
//   1: recip_dir(x) vs 1/x for every float with 2^-60 <= |x| <= 2^60, both signs
//   2: div_ns(a, b) vs a/b on 2^32 hashed pairs inside div_ns's domain
//   3: rcp_ns(x) vs 1/x for every normal float with |x| < 2^126, both signs
//   4: div_ns(a, b) vs a/b for every b in [1, 2) against a of eight mantissas (all ones, all ones
//      but the last bit, 1 ulp above a power of two, a power of two, the half-way mantissas and the
//      alternating patterns) at exponents 0 and 1
__device__ __forceinline__ uint32_t arith_hash(uint64_t k) {
    uint64_t z = k + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return (uint32_t)(z ^ (z >> 31));
}
__global__ void arith_check_kernel(int which, uint64_t n, unsigned long long* bad) {
    unsigned long long nbad = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        bool same;
        if (which == 0) {
            const float x = libm::u2f(0x0F800000u + (uint32_t)i);
            same = libm::f2u(libm::sqrt_ns(x)) == libm::f2u(libm::sqrtf_cr(x));
        } else if (which == 1) {
            const uint32_t span = (uint32_t)(n / 2);
            const uint32_t mag = 0x21800000u + (uint32_t)(i % span);     // 2^-60 .. 2^60
            const float x = libm::u2f(mag | (i >= span ? 0x80000000u : 0u));
            same = libm::f2u(libm::recip_dir(x)) == libm::f2u(1.0f / x);
        } else if (which == 3) {
            const uint32_t span = (uint32_t)(n / 2);
            const float x = libm::u2f((0x00800000u + (uint32_t)(i % span)) | (i >= span ? 0x80000000u : 0u));
            same = libm::f2u(libm::rcp_ns(x)) == libm::f2u(1.0f / x);
        } else if (which == 4) {
            constexpr uint32_t kMant[8] = {0x7FFFFFu, 0x7FFFFEu, 0x000001u, 0x000000u,
                                           0x400000u, 0x3FFFFFu, 0x555555u, 0x2AAAAAu};
            const float b = libm::u2f((127u << 23) | (uint32_t)(i & 0x7FFFFFu));
            const float a = libm::u2f(((127u + (uint32_t)(i >> 26)) << 23) | kMant[(i >> 23) & 7u]);
            same = libm::f2u(libm::div_ns(a, b)) == libm::f2u(a / b);
        } else {
            // |b| in [2^-120, 2^120], |a/b| in [2^-120, 2^120], |a| >= 2^-100, random signs and mantissas
            const uint32_t h0 = arith_hash(2 * i), h1 = arith_hash(2 * i + 1);
            const int eb = (int)(h0 % 241u) - 120;
            int eq = (int)((h0 >> 8) % 241u) - 120;
            int ea = eb + eq;
            if (ea < -100) ea = -100;
            if (ea > 127) ea = 127;
            const float b = libm::u2f(((uint32_t)(eb + 127) << 23) | (h1 & 0x7FFFFFu) | ((h0 >> 31) << 31));
            const float a = libm::u2f(((uint32_t)(ea + 127) << 23) | (arith_hash(i ^ 0x5555u) & 0x7FFFFFu) |
                                      (((h0 >> 30) & 1u) << 31));
            const float q = a / b;
            same = !(fabsf(q) >= 0x1p-120f && fabsf(q) <= 0x1p120f) || libm::f2u(libm::div_ns(a, b)) == libm::f2u(q);
        }
        nbad += same ? 0ull : 1ull;
    }
    if (nbad) atomicAdd(bad, nbad);
}

// Media ids of a caller's uint8 label volume (bits 0-6; bit 7, the detector flag, is dropped)
// packed `bits` per voxel, 32 / bits voxels per word, in the LdsLabels layout. One word per thread.
__global__ void pack_labels_kernel(const uint8_t* lab, uint64_t nvox, uint32_t bits, uint32_t words, uint32_t* out) {
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= words) return;
    const uint32_t per = 32u / bits, mask = (1u << bits) - 1u;
    uint32_t v = 0u;
    for (uint32_t j = 0; j < per; j++) {
        const uint64_t i = (uint64_t)w * per + j;
        if (i < nvox) v |= ((uint32_t)lab[i] & 0x7Fu & mask) << (j * bits);
    }
    out[w] = v;
}

__global__ void libm_eval_kernel(int fn, const float* in, float* out, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float x = in[i], y;
    switch (fn) {
        case 0: y = libm::logf_host(x); break;
        case 1: y = libm::sinf_any(x); break;
        case 2: y = libm::cosf_any(x); break;
        case 3: y = libm::acosf_host(x); break;
        default: y = libm::expf_host(x); break;
    }
    out[i] = y;
}

// ---- deposit-log passes: the log's records summed into the fluence grid ----
// The cells are cut into buckets of 2^kBucketBits consecutive cells. log_hist_kernel counts each
// workgroup's records per bucket, log_scan_kernel turns the counts into each (bucket, workgroup)
// pair's first output slot, log_scatter_kernel writes the records grouped by bucket, and
// log_reduce_kernel sums equal shares of the grouped records in LDS, bucket by bucket, and adds
// the nonzero cells to the grid. hist and scatter take segments s = blockIdx.x, blockIdx.x +
// gridDim.x, ... in the same way, so each workgroup writes exactly the slots it counted.
__device__ __forceinline__ uint32_t log_segments(const unsigned long long* next, uint32_t segs) {
    const unsigned long long n = *next;
    return n < segs ? (uint32_t)n : segs;
}

__global__ void __launch_bounds__(1024) log_hist_kernel(const uint2* log, const uint32_t* fill,
                                                        const unsigned long long* next, uint32_t segs, uint32_t nb,
                                                        uint32_t* hist) {
    extern __shared__ uint32_t lds_u32[];
    uint32_t* h = lds_u32;
    for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) h[b] = 0u;
    __syncthreads();
    const uint32_t nseg = log_segments(next, segs);
    for (uint32_t tile = blockIdx.x; tile * kTileSegs < nseg; tile += gridDim.x)
        for (uint32_t sg = tile * kTileSegs; sg < min((tile + 1) * kTileSegs, nseg); sg++) {
            const uint32_t n = min(fill[sg], kLogSeg);
            const uint2* r = log + (uint64_t)sg * kLogSeg;
            for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) atomicAdd(&h[r[i].x >> kBucketBits], 1u);
        }
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) hist[(uint64_t)b * gridDim.x + blockIdx.x] = h[b];
}

// Exclusive prefix sums over hist (bucket-major: entry b * g + w) in place, one 1024-lane
// workgroup; start[b] = the first slot of bucket b, start[nb] = the record count.
__global__ void __launch_bounds__(1024) log_scan_kernel(uint32_t* hist, uint32_t nb, uint32_t g, uint32_t* start) {
    __shared__ uint32_t part[1024];
    const uint64_t e = (uint64_t)nb * g;
    const uint32_t t = threadIdx.x, nt = blockDim.x;
    const uint64_t lo = e * t / nt, hi = e * (t + 1) / nt;
    uint32_t sum = 0;
    for (uint64_t i = lo; i < hi; i++) sum += hist[i];
    part[t] = sum;
    __syncthreads();
    for (uint32_t o = 1; o < nt; o <<= 1) {
        const uint32_t v = t >= o ? part[t - o] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = part[t] - sum;
    for (uint64_t i = lo; i < hi; i++) {
        const uint32_t c = hist[i];
        hist[i] = run;
        if (i % g == 0) start[i / g] = run;
        run += c;
    }
    if (t == nt - 1) start[nb] = run;
}

// A tile of kTileSegs segments at a time, counting-sorted by bucket in LDS so that each bucket's
// records leave as one run that consecutive lanes write to consecutive slots (scattered 8-byte
// writes had kept the pass at a fifth of its bandwidth): ranks within the tile from LDS counters,
// the buckets' first staged positions from a prefix sum, the records staged in bucket order, then
// written to the workgroup's next slots in their buckets. 1024 lanes, kTileRecs / 1024 records
// each. Tiles are the histogram's: workgroup w takes tiles w, w + gridDim.x, ...
constexpr uint32_t kTileRecs = kTileSegs * kLogSeg;
constexpr uint32_t kPerLane = kTileRecs / 1024;
static_assert(kTileRecs % 1024 == 0, "a tile is a whole number of records per lane");
__global__ void __launch_bounds__(1024) log_scatter_kernel(const uint2* log, const uint32_t* fill,
                                                           const unsigned long long* next, uint32_t segs,
                                                           uint32_t nb, const uint32_t* off, uint2* out) {
    extern __shared__ uint32_t lds_u32[];
    uint32_t* gofs = lds_u32;                          // nb: the workgroup's next slot in each bucket
    uint32_t* cnt = gofs + nb;                         // nb: the tile's records per bucket
    uint32_t* lst = cnt + nb;                          // nb: first staged position of each bucket
    uint2* stage = (uint2*)(lst + nb + (nb & 1u));     // kTileRecs records, 8-byte aligned
    __shared__ uint32_t wsum[16];
    const uint32_t t = threadIdx.x, lane = __lane_id(), wave = t >> 6;
    for (uint32_t b = t; b < nb; b += 1024) gofs[b] = off[(uint64_t)b * gridDim.x + blockIdx.x];
    const uint32_t per = (nb + 1023) / 1024;           // counters per lane in the prefix sum
    const uint32_t b0 = min(t * per, nb), b1 = min(b0 + per, nb);
    const uint32_t nseg = log_segments(next, segs);
    // the records of tile `tile`, record i = t + k * 1024 in v[k]: the first segment's fill first,
    // then the next segments'
    auto load_tile = [&](uint32_t tile, uint2 (&v)[kPerLane]) -> uint32_t {
        uint32_t nfill[kTileSegs];
        uint32_t n = 0;
#pragma unroll
        for (uint32_t j = 0; j < kTileSegs; j++) {
            const uint32_t sg = tile * kTileSegs + j;
            nfill[j] = sg < nseg ? min(fill[sg], kLogSeg) : 0u;
            n += nfill[j];
        }
        const uint2* r = log + (uint64_t)tile * kTileRecs;
#pragma unroll
        for (uint32_t k = 0; k < kPerLane; k++) {
            const uint32_t i = t + k * 1024;
            uint32_t seg = 0, at = i;
#pragma unroll
            for (uint32_t j = 0; j + 1 < kTileSegs; j++)
                if (seg == j && at >= nfill[j]) { at -= nfill[j]; seg = j + 1; }
            v[k] = i < n ? r[(uint64_t)seg * kLogSeg + at] : make_uint2(0u, 0u);
        }
        return n;
    };
    uint2 v[kPerLane];
    uint32_t n = blockIdx.x * kTileSegs < nseg ? load_tile(blockIdx.x, v) : 0u;
    for (uint32_t tile = blockIdx.x; tile * kTileSegs < nseg; tile += gridDim.x) {
        for (uint32_t b = t; b < nb; b += 1024) cnt[b] = 0u;
        __syncthreads();
        uint32_t rk[kPerLane];
#pragma unroll
        for (uint32_t k = 0; k < kPerLane; k++)
            rk[k] = t + k * 1024 < n ? atomicAdd(&cnt[v[k].x >> kBucketBits], 1u) : 0u;
        // the next tile's records load while this one is ranked, staged and written
        uint2 vn[kPerLane];
        const uint32_t nt = tile + gridDim.x;
        uint32_t nn = 0;
        if (nt * kTileSegs < nseg) {
            nn = load_tile(nt, vn);
        } else {
#pragma unroll
            for (uint32_t k = 0; k < kPerLane; k++) vn[k] = make_uint2(0u, 0u);
        }
        __syncthreads();
        uint32_t sum = 0;
        for (uint32_t b = b0; b < b1; b++) sum += cnt[b];
        uint32_t x = sum;                              // inclusive scan over the wave
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o, 64);
            if ((int)lane >= o) x += y;
        }
        if (lane == 63) wsum[wave] = x;
        __syncthreads();
        uint32_t w = lane < 16 ? wsum[lane] : 0u;      // the waves' totals, scanned in every wave
        for (int o = 1; o < 16; o <<= 1) {
            const uint32_t y = __shfl_up(w, o, 64);
            if ((int)lane >= o) w += y;
        }
        const uint32_t before = wave ? (uint32_t)__shfl((int)w, (int)wave - 1, 64) : 0u;
        uint32_t run = before + x - sum;
        for (uint32_t b = b0; b < b1; b++) {
            lst[b] = run;
            run += cnt[b];
        }
        __syncthreads();
#pragma unroll
        for (uint32_t k = 0; k < kPerLane; k++)
            if (t + k * 1024 < n) stage[lst[v[k].x >> kBucketBits] + rk[k]] = v[k];
        __syncthreads();
#pragma unroll
        for (uint32_t k = 0; k < kPerLane; k++) {
            const uint32_t i = t + k * 1024;
            if (i < n) {
                const uint2 q = stage[i];
                const uint32_t b = q.x >> kBucketBits;
                out[gofs[b] + (i - lst[b])] = q;
            }
        }
        __syncthreads();
        for (uint32_t b = t; b < nb; b += 1024) gofs[b] += cnt[b];
#pragma unroll
        for (uint32_t k = 0; k < kPerLane; k++) v[k] = vn[k];
        n = nn;
    }
}

// Adds the values of N records {cell, value bits} to their cells of a 2^kBucketBits-cell LDS tile
// of floats, by compare-and-swap: the N current values are read, the N sums formed in registers
// and swapped in together, and only a swap that found its cell changed meanwhile (by another lane,
// or by an earlier record of the same lane) is formed again from the value it found. Each float
// add is rounded as an atomic add's; on gfx950 ds_add_f32 took 3.3x as long in the reduce pass
// (8.98 against 2.73 ms per 1.9e9 records; the integer swap runs at the pass's load speed).
template <int N>
__device__ __forceinline__ void lds_add_records(uint32_t* tile, const uint2 (&v)[N]) {
    constexpr uint32_t kMask = (1u << kBucketBits) - 1u;
    uint32_t old[N], prev[N];
#pragma unroll
    for (int u = 0; u < N; u++)
        old[u] = __hip_atomic_load(tile + (v[u].x & kMask), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
#pragma unroll
    for (int u = 0; u < N; u++)
        prev[u] = atomicCAS(tile + (v[u].x & kMask), old[u], __float_as_uint(__uint_as_float(old[u]) + __uint_as_float(v[u].y)));
#pragma unroll
    for (int u = 0; u < N; u++)
        while (prev[u] != old[u]) {   // ends: each failed swap means another swap on the cell succeeded
            old[u] = prev[u];
            prev[u] = atomicCAS(tile + (v[u].x & kMask), old[u],
                                __float_as_uint(__uint_as_float(old[u]) + __uint_as_float(v[u].y)));
        }
}

// Workgroup w of gridDim.x sums the grouped records [total * w / gridDim.x, total * (w + 1) /
// gridDim.x): equal shares of records, whatever the buckets hold (a few buckets near a source can
// hold most of the records). The share's records of one bucket are summed in a 2^kBucketBits-cell
// LDS tile, whose nonzero cells are then added to the grid and cleared before the share's next
// bucket; a bucket split between workgroups is added by each of them.
__global__ void __launch_bounds__(1024) log_reduce_kernel(const uint2* sorted, const uint32_t* start, uint32_t nb,
                                                          uint64_t ncells, float* grid) {
    extern __shared__ float lds_f32[];
    float* tile = lds_f32;
    constexpr uint32_t kCells = 1u << kBucketBits;
    const uint32_t total = start[nb];
    const uint32_t lo = (uint32_t)((uint64_t)total * blockIdx.x / gridDim.x);
    const uint32_t hi = (uint32_t)((uint64_t)total * (blockIdx.x + 1) / gridDim.x);
    if (lo == hi) return;                    // the whole workgroup: no barrier is skipped by part of it
    for (uint32_t i = threadIdx.x; i < kCells; i += blockDim.x) tile[i] = 0.f;
    __syncthreads();
    for (uint32_t pos = lo; pos < hi;) {
        // every lane reads the same record: its bucket, and where the bucket's records end
        const uint32_t b = (uint32_t)__builtin_amdgcn_readfirstlane((int)(sorted[pos].x >> kBucketBits));
        uint32_t end = b < nb ? (uint32_t)__builtin_amdgcn_readfirstlane((int)start[b + 1]) : hi;
        if (end <= pos || end > hi) end = hi;   // each round advances, whatever the records hold
        // eight records per lane in flight before their LDS adds
        constexpr uint32_t kUnroll = 8;
        uint32_t i = pos + threadIdx.x;
        for (; i + (kUnroll - 1) * blockDim.x < end; i += kUnroll * blockDim.x) {
            uint2 v[kUnroll];
#pragma unroll
            for (uint32_t u = 0; u < kUnroll; u++) v[u] = sorted[i + u * blockDim.x];
            lds_add_records<kUnroll>((uint32_t*)tile, v);
        }
        for (; i < end; i += blockDim.x) {
            uint2 v[1] = {sorted[i]};
            lds_add_records<1>((uint32_t*)tile, v);
        }
        __syncthreads();
        const uint64_t base = (uint64_t)b << kBucketBits;
        for (uint32_t c = threadIdx.x; c < kCells; c += blockDim.x) {
            const float v = tile[c];
            if (v != 0.f) {
                if (base + c < ncells) unsafeAtomicAdd(grid + base + c, v);
                tile[c] = 0.f;
            }
        }
        __syncthreads();
        pos = end;
    }
}

}  // namespace dev

// ======================= host side =======================

namespace {

using LaunchFn = void (*)(const dev::KernelArgs&, int, int, size_t, hipStream_t);

template <bool R, bool D, typename L, bool P = false, int S = dev::kLabHbm, bool G = false>
void launch_one(const dev::KernelArgs& ka, int blocks, int block, size_t lds, hipStream_t s) {
    hipLaunchKernelGGL((dev::photon_kernel<R, D, L, P, S, G>), dim3(blocks), dim3(block), lds, s, ka);
}

// kernel variants: reflection, detectors, label type, replay, label mode (LDS-staged ids with
// uint8 labels only), deposit log (G)
template <bool P, bool G, int S>
LaunchFn pick_mode(bool reflect, bool savedet, bool u8) {
    if (u8 || S != dev::kLabHbm) {
        if (reflect)
            return savedet ? launch_one<true, true, uint8_t, P, S, G> : launch_one<true, false, uint8_t, P, S, G>;
        return savedet ? launch_one<false, true, uint8_t, P, S, G> : launch_one<false, false, uint8_t, P, S, G>;
    }
    if (reflect)
        return savedet ? launch_one<true, true, uint32_t, P, S, G> : launch_one<true, false, uint32_t, P, S, G>;
    return savedet ? launch_one<false, true, uint32_t, P, S, G> : launch_one<false, false, uint32_t, P, S, G>;
}
template <bool P, bool G = false>
LaunchFn pick_variant(bool reflect, bool savedet, bool u8, int lab) {
    if (lab == dev::kLabLds) return pick_mode<P, G, dev::kLabLds>(reflect, savedet, u8);
    return pick_mode<P, G, dev::kLabHbm>(reflect, savedet, u8);
}

LaunchFn pick_replay(bool reflect, bool u8, int lab) { return pick_variant<true>(reflect, false, u8, lab); }

LaunchFn pick(bool reflect, bool savedet, bool u8, int lab, bool dlog = false) {
    return dlog ? pick_variant<false, true>(reflect, savedet, u8, lab) : pick_variant<false>(reflect, savedet, u8, lab);
}

template <bool R, bool D, typename L, int S>
const void* kptr() {
    return (const void*)dev::photon_kernel<R, D, L, false, S>;
}
template <int S>
const void* kptr_mode(bool reflect, bool savedet, bool u8) {
    if (u8 || S != dev::kLabHbm) {
        if (reflect) return savedet ? kptr<true, true, uint8_t, S>() : kptr<true, false, uint8_t, S>();
        return savedet ? kptr<false, true, uint8_t, S>() : kptr<false, false, uint8_t, S>();
    }
    if (reflect) return savedet ? kptr<true, true, uint32_t, S>() : kptr<true, false, uint32_t, S>();
    return savedet ? kptr<false, true, uint32_t, S>() : kptr<false, false, uint32_t, S>();
}
const void* kernel_ptr(bool reflect, bool savedet, bool u8, int lab) {
    if (lab == dev::kLabLds) return kptr_mode<dev::kLabLds>(reflect, savedet, u8);
    return kptr_mode<dev::kLabHbm>(reflect, savedet, u8);
}

template <typename T>
T* dalloc(size_t n) {
    void* p = nullptr;
    if (n == 0) n = 1;
    HIP_OK(hipMalloc(&p, n * sizeof(T)));
    return (T*)p;
}

}  // namespace

int device_count() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

std::string device_name(int device) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return "";
    return std::string(prop.name) + " (" + prop.gcnArchName + ")";
}

Session::Session(const Problem& p, const LaunchConfig& lc) : prob_(p), lc_(lc) {
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

    const uint64_t nvox = (uint64_t)p.nx * p.ny * p.nz;
    label_bytes_ = p.mediumnum <= 128 ? 1 : 4;
    if (label_bytes_ == 1) {
        std::vector<uint8_t> lab(nvox);
        for (uint64_t i = 0; i < nvox; i++) {
            uint32_t v = (uint32_t)p.labels[i];
            lab[i] = (uint8_t)((v & kMedMask) | ((v & kDetMask) ? 0x80u : 0u));
        }
        d_labels_ = dalloc<uint8_t>(nvox);
        HIP_OK(hipMemcpy(d_labels_, lab.data(), nvox, hipMemcpyHostToDevice));
        // media ids packed for LDS staging, when the whole volume fits: 2 or 4 bits holding the id
        // (up to 4 or 16 media); a volume holding at most two media ids also gets a 1-bit form
        // (code 0 = the smaller id), which leaves more LDS to the tally tile and is switched to
        // after the pilot when most deposits would otherwise go to HBM (or used from the start
        // when it is the only form that fits)
        uint32_t ids[2] = {~0u, ~0u};
        int nids = 0;
        for (uint64_t i = 0; i < nvox && nids <= 2; i++) {
            const uint32_t m = (uint32_t)p.labels[i] & kMedMask;
            if (nids > 0 && m == ids[0]) continue;
            if (nids > 1 && m == ids[1]) continue;
            if (nids < 2) ids[nids] = m;
            nids++;
        }
        if (nids == 2 && ids[1] < ids[0]) std::swap(ids[0], ids[1]);
        if (nids == 1) ids[1] = ids[0];
        // The media ids the volume uses: when they form an arithmetic run id0, id0 + step, ...
        // of at most 16, the codes (id - id0) / step take 2 bits (4 ids) or 4 bits (16 ids)
        // whatever the ids themselves are (the 200^3 skin model uses ids 1-4 of 5 media: 2 bits);
        // otherwise the ids are packed as they are, for volumes of at most 16 media.
        std::vector<uint8_t> seen(128, 0);
        for (uint64_t i = 0; i < nvox; i++) seen[(uint32_t)p.labels[i] & 0x7Fu] = 1;
        std::vector<uint32_t> used;
        for (uint32_t m = 0; m < 128; m++)
            if (seen[m]) used.push_back(m);
        uint32_t cid0 = 0, cstep = 1, wbits = p.mediumnum <= 4 ? 2u : (p.mediumnum <= 16 ? 4u : 0u);
        if (!used.empty() && used.size() <= 16) {
            const uint32_t d = used.size() > 1 ? used[1] - used[0] : 1u;
            bool run = true;
            for (size_t k = 0; k < used.size(); k++) run = run && used[k] == used[0] + (uint32_t)k * d;
            if (run) {
                cid0 = used[0];
                cstep = d;
                wbits = used.size() <= 4 ? 2u : 4u;
            }
        }
        auto pack = [&](uint32_t bits, uint32_t id0, uint32_t step, uint64_t* words_out,
                        uint64_t max_bytes) -> uint32_t* {
            const uint64_t words = (nvox * bits + 31) / 32;
            if (lc.label_cache < 0 || words * 4 > max_bytes) return nullptr;
            std::vector<uint32_t> packed(words, 0u);
            const uint32_t per = 32 / bits;
            for (uint64_t i = 0; i < nvox; i++) {
                const uint32_t code = (((uint32_t)p.labels[i] & kMedMask) - id0) / step;
                packed[i / per] |= code << ((i % per) * bits);
            }
            uint32_t* d = dalloc<uint32_t>(words);
            HIP_OK(hipMemcpy(d, packed.data(), words * 4, hipMemcpyHostToDevice));
            *words_out = words;
            return d;
        };
        uint64_t words = 0;
        if (wbits && (d_plabels_ = pack(wbits, cid0, cstep, &words, kLabelCacheBytes))) {
            plabel_words_ = (uint32_t)words;
            plabel_bits_ = wbits;
            plabel_id0_ = cid0;
            plabel_step_ = cstep;
        }
        if (nids <= 2) {
            const uint32_t step1 = ids[1] > ids[0] ? ids[1] - ids[0] : 1u;
            uint32_t* d1 = pack(1, ids[0], step1, &words, kLabelCacheBytes);
            if (d1 && !d_plabels_) {                 // the 1-bit form is the only one that fits
                d_plabels_ = d1;
                plabel_words_ = (uint32_t)words;
                plabel_bits_ = 1;
                plabel_id0_ = ids[0];
                plabel_step_ = step1;
            } else if (d1) {
                d_plabels1_ = d1;
                plabel1_words_ = (uint32_t)words;
                plabel1_id0_ = ids[0];
                plabel1_step_ = step1;
            }
        }
    } else {
        d_labels_ = dalloc<uint32_t>(nvox);
        HIP_OK(hipMemcpy(d_labels_, p.labels.data(), nvox * 4, hipMemcpyHostToDevice));
    }
    d_media_ = dalloc<float>(p.media.size());
    HIP_OK(hipMemcpy(d_media_, p.media.data(), p.media.size() * 4, hipMemcpyHostToDevice));
    d_detpos_ = dalloc<float>(std::max<size_t>(p.detpos.size(), 4));
    if (!p.detpos.empty()) HIP_OK(hipMemcpy(d_detpos_, p.detpos.data(), p.detpos.size() * 4, hipMemcpyHostToDevice));
    fluence_len_ = p.issavevol ? nvox * (uint64_t)p.maxgate : 0;
    d_fluence_ = dalloc<float>(fluence_len_);
    detcap_ = p.issavedet ? p.maxdetphotons : 0;
    int cols = std::max(p.detp_columns(), 1);
    d_detp_ = dalloc<float>((size_t)detcap_ * cols);
    d_detid_ = dalloc<uint32_t>(detcap_);
    d_detw_ = dalloc<float>(detcap_);
    d_counters_ = dalloc<uint32_t>(4);
    d_cursor_ = dalloc<unsigned long long>(2);
    d_energy_ = dalloc<double>(4);

    // launch shape: fill every CU with persistent waves
    hipDeviceProp_t prop;
    HIP_OK(hipGetDeviceProperties(&prop, lc.device));
    // with the labels in LDS, one 1024-lane workgroup per CU holds one copy of them
    block_ = lc.block > 0 ? lc.block : (plabel_words_ ? 1024 : (p.blocksize > 0 ? p.blocksize : 256));
    block_ = std::max(64, std::min(1024, (block_ / 64) * 64));
    size_t lds = lds_bytes(0);
    const size_t lds_wg_max = (prop.sharedMemPerBlock ? prop.sharedMemPerBlock : 64 * 1024) - kStaticLds;
    while (lds > lds_wg_max && block_ > 64) {
        block_ /= 2;
        lds = lds_bytes(0);
    }
    if (lds > lds_wg_max) throw std::runtime_error("too many media for the LDS staging of media and partial paths");
    int per_cu = 0;
    const bool reflect = p.isreflect != 0, savedet = p.issavedet != 0;
    HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel_ptr(reflect, savedet, label_bytes_ == 1, label_mode()), block_, lds));
    if (per_cu < 1) per_cu = 1;
    // labels in HBM: at most kHbmLabelWgPerCu workgroups per CU, so each keeps a 40 KiB LDS share
    // for its tally tile (200^3 skin model: 2256 ms at 4 per CU, 2304 ms at the 5 registers allow)
    if (!plabel_words_) per_cu = std::min(per_cu, kHbmLabelWgPerCu);
    // LDS tally tile: share the CU's LDS between the workgroups the register budget lets in
    tally_mode_ = lc.tally_mode;
    const size_t lds_cu = prop.maxSharedMemoryPerMultiProcessor ? prop.maxSharedMemoryPerMultiProcessor : 160 * 1024;
    per_wg_lds_ = std::min(lds_cu / (size_t)per_cu - kStaticLds, lds_wg_max);
    tile_budget_ = per_wg_lds_ > lds ? std::min<uint64_t>((per_wg_lds_ - lds) / 4, kMaxTileCells) : 0;
    if (tally_mode_ == 1 && tile_budget_ < 64) tile_budget_ = 64;
    blocks_ = lc.blocks > 0 ? lc.blocks : prop.multiProcessorCount * per_cu;
    if (p.threadnum > 0 && lc.blocks <= 0) blocks_ = std::max(1, p.threadnum / block_);
    cu_count_ = std::max(1, prop.multiProcessorCount);
    auto_shape_ = lc.block <= 0 && lc.blocks <= 0 && p.threadnum <= 0 && p.blocksize <= 0;
    reset();
}

Session::~Session() {
    (void)hipSetDevice(lc_.device);
    for (void* ptr : {(void*)d_labels_, (void*)d_media_, (void*)d_detpos_, (void*)d_fluence_, (void*)d_detp_,
                      (void*)d_detid_, (void*)d_detw_, (void*)d_counters_, (void*)d_cursor_, (void*)d_energy_,
                      (void*)d_trace_, d_vprops_, (void*)d_plabels_, (void*)d_plabels1_, (void*)d_msrc_, d_dlog_,
                      d_dlog_sorted_, (void*)d_dlog_fill_, (void*)d_dlog_next_, (void*)d_dlog_hist_,
                      (void*)d_dlog_start_})
        if (ptr) (void)hipFree(ptr);
    if (ev0_) (void)hipEventDestroy((hipEvent_t)ev0_);
    if (ev1_) (void)hipEventDestroy((hipEvent_t)ev1_);
    if (stream_) (void)hipStreamDestroy((hipStream_t)stream_);
}

void Session::reset() {
    HIP_OK(hipSetDevice(lc_.device));
    hipStream_t s = (hipStream_t)stream_;
    if (fluence_len_) HIP_OK(hipMemsetAsync(d_fluence_, 0, fluence_len_ * 4, s));
    HIP_OK(hipMemsetAsync(d_counters_, 0, 16, s));
    HIP_OK(hipMemsetAsync(d_energy_, 0, 32, s));
    HIP_OK(hipStreamSynchronize(s));
}

void* Session::build_args(uint64_t first, uint64_t count, void* out) const {
    const Problem& p = prob_;
    dev::KernelArgs& ka = *(dev::KernelArgs*)out;
    memset(&ka, 0, sizeof ka);
    ka.nx = p.nx; ka.ny = p.ny; ka.nz = p.nz; ka.nyz = p.ny * p.nz;
    ka.nxyz = (uint64_t)p.nx * p.ny * p.nz;
    ka.maxgate = p.maxgate; ka.outputtype = p.outputtype; ka.savedetflag = p.savedetflag;
    ka.mediumnum = p.mediumnum; ka.detnum = p.issavedet ? p.detnum : 0; ka.srctype = p.srctype;
    ka.issavevol = p.issavevol; ka.ppathlen = p.ppath_columns(); ka.detlen = p.detp_columns();
    ka.tstart = p.tstart; ka.tend = p.tend; ka.rtstep = p.rtstep;
    ka.srcpos = make_float4(p.srcpos[0], p.srcpos[1], p.srcpos[2], p.srcpos[3]);
    ka.srcdir = make_float4(p.srcdir[0], p.srcdir[1], p.srcdir[2], p.srcdir[3]);
    ka.srcp1 = make_float4(p.srcparam1[0], p.srcparam1[1], p.srcparam1[2], p.srcparam1[3]);
    ka.srcp2 = make_float4(p.srcparam2[0], p.srcparam2[1], p.srcparam2[2], p.srcparam2[3]);
    for (int i = 0; i < 4; i++) ka.seeds[i] = p.seeds[i];
    ka.first = first; ka.count = count;
    ka.labels = d_labels_;
    ka.media = (const float4*)d_media_;
    ka.detpos = (const float4*)d_detpos_;
    ka.fluence = d_fluence_;
    ka.detp = d_detp_; ka.detid = d_detid_; ka.detw = d_detw_; ka.ndet = d_counters_; ka.maxdet = detcap_;
    ka.energy = d_energy_;
    ka.cursor = d_cursor_;
    ka.trace = lc_.trace ? d_trace_ : nullptr;
    ka.vote = vote();
    ka.dep_cache = deposit_cache() ? 1 : 0;
    ka.vprops = (const float2*)d_vprops_;
    ka.unit = p.unitinmm;
    ka.plabels = d_plabels_;
    ka.plabel_words = plabel_words_;
    ka.plabel_bits = plabel_bits_;
    ka.plabel_lg = plabel_bits_ == 1 ? 5u : (plabel_bits_ == 2 ? 4u : 3u);
    ka.plabel_id0 = plabel_id0_;
    ka.plabel_step = plabel_step_;
    ka.per_wg = (count - count / 8) / (uint64_t)std::max(blocks_, 1);
    ka.msrc_n = msrc_n_;
    ka.rsrcs = msrc_n_ ? d_msrc_ : nullptr;
    if (tile_[4] > 0) {
        ka.tx0 = tile_[0]; ka.ty0 = tile_[1]; ka.tz0 = tile_[2]; ka.tg0 = tile_[3];
        ka.tsx = tile_[4]; ka.tsy = tile_[5]; ka.tsz = tile_[6]; ka.tsg = tile_[7];
        ka.tcells = ka.tsx * ka.tsy * ka.tsz;
        ka.tile_n = ka.tcells * ka.tsg;
    }
    return out;
}

size_t Session::lds_bytes(uint32_t tile_n, bool paths_needed) const {
    const size_t paths = paths_needed && prob_.ppath_columns() > kRegPaths ? (size_t)prob_.ppath_columns() * block_ * 4 : 0;
    return 32 * 16 + (size_t)prob_.mediumnum * 16 + (size_t)plabel_words_ * 4 + (size_t)tile_n * 4 + paths;
}

// The tally-box search on a count grid in device memory ([t][x][y][z], n cells): the densest
// cell, the total, then the greedy growth in one workgroup. scratch: 8 device words (densest
// cell, largest count bits, total, counts in the box, the box as 8 ints), cleared here.
// Enqueued on s; read scratch back after the stream is synchronised.
static void enqueue_box_search(const float* d_cnt, uint64_t n, int nt, int nx, int ny, int nz, uint64_t budget,
                               unsigned long long* scratch, hipStream_t s) {
    HIP_OK(hipMemsetAsync(scratch, 0, 64, s));
    HIP_OK(hipMemsetAsync(scratch, 0xFF, 8, s));                       // densest cell starts at ~0
    dev::count_max_kernel<<<2048, 256, 0, s>>>(d_cnt, n, (unsigned*)(scratch + 1));
    dev::count_argmax_sum_kernel<<<2048, 256, 0, s>>>(d_cnt, n, (const unsigned*)(scratch + 1), scratch,
                                                      (double*)(scratch + 2));
    dev::grow_box_kernel<<<1, 1024, 0, s>>>(d_cnt, nt, nx, ny, nz, budget, scratch, (int*)(scratch + 4),
                                            (double*)(scratch + 3));
    HIP_OK(hipGetLastError());
}

// Pick the LDS tally tile from where a pilot sample of this photon range deposits most often.
// The count grid never leaves the device: two passes find the densest cell (the first of equal
// ones) and the total, and one workgroup grows the box from that cell exactly as choose_box
// does on the host. The pilot's deposits go through the per-lane cache only where its 32-bit
// keys address every cell.
void Session::choose_tile(uint64_t first, uint64_t count) {
    tile_chosen_ = true;
    const Problem& p = prob_;
    if (!p.issavevol || tally_mode_ == 0 || tile_budget_ < 64) return;
    hipStream_t s = (hipStream_t)stream_;
    uint64_t pilot = std::min<uint64_t>(count, 262144);
    const uint64_t n = fluence_len_;
    const bool cache = n < 0xFFFFFFFFull;
    float* d_cnt = dalloc<float>(n);
    double* d_e = dalloc<double>(4);
    // 0-1 cursor, 2-4 deposit and scatter statistics, 8-15 the box search
    unsigned long long* d_c = dalloc<unsigned long long>(16);
    HIP_OK(hipMemsetAsync(d_cnt, 0, n * 4, s));
    HIP_OK(hipMemsetAsync(d_e, 0, 32, s));
    HIP_OK(hipMemsetAsync(d_c, 0, 64, s));
    dev::KernelArgs ka;
    build_args(first, pilot, &ka);
    ka.fluence = d_cnt; ka.energy = d_e; ka.cursor = d_c; ka.trace = nullptr;
    ka.detnum = 0; ka.ppathlen = 0; ka.count_mode = 1;
    ka.dep_cache = cache ? 1 : 0; ka.dep_stats = cache ? d_c + 2 : nullptr;
    ka.tile_n = 0; ka.tsx = ka.tsy = ka.tsz = ka.tsg = 0;
    const bool reflect = p.isreflect != 0;
    pick(reflect, false, label_bytes_ == 1, label_mode())(ka, blocks_, block_, lds_bytes(0, false), s);
    HIP_OK(hipGetLastError());
    enqueue_box_search(d_cnt, n, p.maxgate, (int)p.nx, (int)p.ny, (int)p.nz, tile_budget_, d_c + 8, s);
    unsigned long long head[16];
    HIP_OK(hipMemcpyAsync(head, d_c, sizeof head, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    dep_hits_ = head[2] ? 1.0 - (double)head[3] / (double)head[2] : 0.0;
    double total, in_box;
    memcpy(&total, &head[10], 8);
    memcpy(&in_box, &head[11], 8);
    memcpy(tile_, &head[12], sizeof(int) * 8);
    // most deposits outside the tile: where the volume has a 1-bit label form, take it and the
    // larger tile the LDS it frees allows (no change to any photon, only to where it tallies)
    if (d_plabels1_ && (total > 0 ? in_box / total : 0.0) < kCacheBelowShare) {
        const size_t lds1 = lds_bytes(0) - (size_t)plabel_words_ * 4 + (size_t)plabel1_words_ * 4;
        const uint64_t budget1 =
            per_wg_lds_ > lds1 ? std::min<uint64_t>((per_wg_lds_ - lds1) / 4, kMaxTileCells) : 0;
        if (budget1 > tile_budget_) {
            enqueue_box_search(d_cnt, n, p.maxgate, (int)p.nx, (int)p.ny, (int)p.nz, budget1, d_c + 8, s);
            HIP_OK(hipMemcpyAsync(head, d_c, sizeof head, hipMemcpyDeviceToHost, s));
            HIP_OK(hipStreamSynchronize(s));
            memcpy(&in_box, &head[11], 8);
            memcpy(tile_, &head[12], sizeof(int) * 8);
            std::swap(d_plabels_, d_plabels1_);
            std::swap(plabel_words_, plabel1_words_);
            plabel_bits_ = 1;
            plabel_id0_ = plabel1_id0_;
            plabel_step_ = plabel1_step_;
            tile_budget_ = budget1;
        }
    }
    (void)hipFree(d_cnt); (void)hipFree(d_e); (void)hipFree(d_c);
    // the per-lane deposit cache pays where most deposits go to HBM and a fair share of them hit
    // it (the pilot ran with it); otherwise its extra work per deposit costs more than it saves
    dep_cache_ = (total > 0 ? in_box / total : 0.0) < kCacheBelowShare && dep_hits_ >= kCacheMinHits;
    // global deposits per photon the runs will issue (outside the tile, after the cache): past
    // kLogMinRecords they would saturate the memory-side atomics, and the deposit log takes them
    const double share_out = total > 0 ? 1.0 - in_box / total : 1.0;
    dlog_rpp_ = cache ? (double)head[2] / (double)pilot * share_out * (deposit_cache() ? 1.0 - dep_hits_ : 1.0) : 0.0;
    dlog_auto_ = dlog_rpp_ >= kLogMinRecords;
    // few scatters per voxel crossing (small voxels, or media that scatter little): the scatter
    // code runs with a few lanes of each wave in most iterations, and gating it on a lane vote
    // pays (200^3 skin model at 0.11 scatters per crossing: 6 % faster at 16 lanes; the
    // heterogeneous reflective cube at 0.83: 2-9 % slower at any threshold)
    vote_auto_ = (cache && head[2] > 0 && (double)head[4] / (double)head[2] < kVoteBelowScatters) ? kVoteLanes : 0;
}

bool Session::deposit_log() const {
    if (!prob_.issavevol || dlog_unavailable_ || fluence_len_ > ((uint64_t)kMaxBuckets << kBucketBits)) return false;
    return lc_.deposit_log > 0 || (lc_.deposit_log == 0 && dlog_auto_);
}

// Log buffers for a run of `count` photons: as many records as the pilot expects (times 1.25),
// at least four partly filled segments' worth per wave, at most kLogMaxRecords and a quarter of the
// free device memory for the log and its grouped copy together; runs that need more are split
// into chunks. An explicit capacity (LaunchConfig::deposit_log_segments) is taken as given.
void Session::ensure_log(uint64_t count) {
    const double rpp = dlog_rpp_ > 0.0 ? dlog_rpp_ : 512.0;
    const uint64_t waves = (uint64_t)blocks_ * (uint64_t)((block_ + 63) / 64);
    uint64_t want;
    if (lc_.deposit_log_segments > 0) {
        want = (uint64_t)lc_.deposit_log_segments * kLogSeg;
    } else {
        size_t free_b = 0, total_b = 0;
        HIP_OK(hipMemGetInfo(&free_b, &total_b));
        const uint64_t cap = std::min<uint64_t>(kLogMaxRecords, (uint64_t)(free_b + 16 * dlog_records_) / 4 / 16);
        want = std::max<uint64_t>((uint64_t)((double)count * rpp * 1.25), 4 * waves * kLogSeg);
        want = std::min(want, cap) / kLogSeg * kLogSeg;
        if (want < kLogSeg) throw std::runtime_error("no device memory left for the deposit log");
    }
    if (dlog_records_ >= want) return;
    for (void* ptr : {d_dlog_, d_dlog_sorted_, (void*)d_dlog_fill_, (void*)d_dlog_next_, (void*)d_dlog_hist_,
                      (void*)d_dlog_start_})
        if (ptr) HIP_OK(hipFree(ptr));
    const uint32_t nb = (uint32_t)((fluence_len_ + (1ull << kBucketBits) - 1) >> kBucketBits);
    d_dlog_ = dalloc<uint2>(want);
    d_dlog_sorted_ = dalloc<uint2>(want);
    d_dlog_fill_ = dalloc<uint32_t>(want / kLogSeg);
    d_dlog_next_ = dalloc<unsigned long long>(1);
    d_dlog_hist_ = dalloc<uint32_t>((size_t)nb * kLogGridPerCu * cu_count_);
    d_dlog_start_ = dalloc<uint32_t>(nb + 1);
    dlog_records_ = want;
}

// The log for a run, or false when the device cannot hold it (the run then uses float atomics,
// with the same results up to summation order).
bool Session::log_ready(uint64_t count) {
    try {
        ensure_log(count);
        return true;
    } catch (const std::exception&) {
        for (void* ptr : {d_dlog_, d_dlog_sorted_, (void*)d_dlog_fill_, (void*)d_dlog_next_, (void*)d_dlog_hist_,
                          (void*)d_dlog_start_})
            if (ptr) (void)hipFree(ptr);
        d_dlog_ = d_dlog_sorted_ = nullptr;
        d_dlog_fill_ = d_dlog_hist_ = d_dlog_start_ = nullptr;
        d_dlog_next_ = nullptr;
        dlog_records_ = 0;
        dlog_unavailable_ = true;  // the session stays on float atomics
        (void)hipGetLastError();   // an allocation failure leaves no lasting error state
        return false;
    }
}

// The passes that add the log's records to the fluence grid, enqueued after a photon launch
// expected to log about expect_records records.
void Session::enqueue_log_passes(void* stream, uint64_t expect_records) {
    hipStream_t s = (hipStream_t)stream;
    const uint32_t nb = (uint32_t)((fluence_len_ + (1ull << kBucketBits) - 1) >> kBucketBits);
    const uint32_t segs = (uint32_t)(dlog_records_ / kLogSeg);
    const int g = kLogGridPerCu * cu_count_;
    const size_t hist_lds = (size_t)nb * 4;
    hipLaunchKernelGGL(dev::log_hist_kernel, dim3(g), dim3(1024), hist_lds, s, (const uint2*)d_dlog_, d_dlog_fill_,
                       d_dlog_next_, segs, nb, d_dlog_hist_);
    hipLaunchKernelGGL(dev::log_scan_kernel, dim3(1), dim3(1024), 0, s, d_dlog_hist_, nb, (uint32_t)g, d_dlog_start_);
    const size_t scatter_lds = ((size_t)3 * nb + (nb & 1u)) * 4 + (size_t)kTileSegs * kLogSeg * 8;
    hipLaunchKernelGGL(dev::log_scatter_kernel, dim3(g), dim3(1024), scatter_lds, s, (const uint2*)d_dlog_, d_dlog_fill_,
                       d_dlog_next_, segs, nb, d_dlog_hist_, (uint2*)d_dlog_sorted_);
    // equal shares of about kReduceShare records, from one to kReduceGridPerCu workgroups per CU
    const uint64_t shares = std::min<uint64_t>(std::max<uint64_t>(expect_records / kReduceShare, cu_count_),
                                               (uint64_t)kReduceGridPerCu * cu_count_);
    hipLaunchKernelGGL(dev::log_reduce_kernel, dim3((uint32_t)shares), dim3(1024), (size_t)4 << kBucketBits, s,
                       (const uint2*)d_dlog_sorted_, d_dlog_start_, nb, fluence_len_, d_fluence_);
    HIP_OK(hipGetLastError());
}

// Launch shape of one run. The full shape fills every CU; a run of fewer than
// kShortRunPhotonsPerLane photons per lane is bound by its slowest photons, whose waves issue
// more often when fewer waves share a SIMD, so it runs on half the lanes: half-size workgroups
// when one workgroup per CU holds the staged labels (same LDS layout and tally tile), else half
// the workgroups while at least two per CU remain. Explicit shapes are kept as given.
void Session::run_shape(uint64_t count, int* block, int* blocks) const {
    *block = block_;
    *blocks = blocks_;
    if (!auto_shape_ || count >= kShortRunPhotonsPerLane * (uint64_t)block_ * (uint64_t)blocks_) return;
    if (!plabel_words_ && blocks_ >= 2 * cu_count_) *blocks = blocks_ / 2;
    else if (block_ >= 128) *block = (block_ / 2 / 64) * 64;
}

RunStats Session::run(uint64_t first, uint64_t count) {
    HIP_OK(hipSetDevice(lc_.device));
    hipStream_t s = (hipStream_t)stream_;
    const Problem& p = prob_;
    RunStats st;
    st.photons = count;
    // with several sources, slot p reads source p / msrc_n_ of the table: a range past the last
    // source would read beyond it
    if (msrc_n_) {
        const uint64_t total = msrc_count_ * msrc_n_;
        if (first > total || count > total - first)
            throw std::runtime_error("umcx_session_run: photons [" + std::to_string(first) + ", " +
                                     std::to_string(first + count) + ") exceed the " + std::to_string(msrc_count_) +
                                     " sources x " + std::to_string(msrc_n_) + " photons set by umcx_session_set_sources");
    }
    last_count_ = count;
    if (lc_.trace) {
        if (trace_cap_ < count) {
            if (d_trace_) HIP_OK(hipFree(d_trace_));
            d_trace_ = dalloc<TraceRec>(count);
            trace_cap_ = count;
        }
        HIP_OK(hipMemsetAsync(d_trace_, 0, count * sizeof(TraceRec), s));
    }
    if (count == 0) return st;
    if (!tile_chosen_ && (tally_mode_ == 1 || (tally_mode_ < 0 && count >= 1000000))) choose_tile(first, count);
    dev::KernelArgs ka;
    build_args(first, count, &ka);
    int rblock = block_, rblocks = blocks_;
    run_shape(count, &rblock, &rblocks);
    ka.per_wg = (count - count / 8) / (uint64_t)rblocks;
    last_block_ = rblock;
    last_blocks_ = rblocks;
    const bool reflect = p.isreflect != 0, savedet = p.issavedet != 0;
    size_t lds = lds_bytes(ka.tile_n);
    const LaunchFn launch = pick(reflect, savedet, label_bytes_ == 1, label_mode());
    if (deposit_log() && log_ready(count)) {
        const LaunchFn launch_log = pick(reflect, savedet, label_bytes_ == 1, label_mode(), true);
        // global deposits go to the log; a run larger than the log holds runs in chunks, each
        // launch followed by the passes that add its records to the grid
        const double rpp = dlog_rpp_ > 0.0 ? dlog_rpp_ : 512.0;
        const uint64_t waves = (uint64_t)rblocks * (uint64_t)((rblock + 63) / 64);
        // a wave may leave up to two segments partly filled
        const uint64_t room = dlog_records_ > 2 * waves * kLogSeg ? dlog_records_ - 2 * waves * kLogSeg : 0;
        uint64_t per_chunk = std::max<uint64_t>(1, (uint64_t)((double)room / (rpp * 1.1)));
        // an explicit capacity: at most four chunks (deposits past a full log go to atomics)
        if (lc_.deposit_log_segments > 0) per_chunk = std::max<uint64_t>(per_chunk, (count + 3) / 4);
        HIP_OK(hipEventRecord((hipEvent_t)ev0_, s));
        for (uint64_t done = 0; done < count;) {
            const uint64_t n = std::min(per_chunk, count - done);
            dev::KernelArgs kc = ka;
            kc.first = first + done;
            kc.count = n;
            int cblock = block_, cblocks = blocks_;
            run_shape(n, &cblock, &cblocks);
            kc.per_wg = (n - n / 8) / (uint64_t)cblocks;
            if (kc.trace) kc.trace = ka.trace + done;   // trace entries are per photon of the run
            kc.dlog = (uint2*)d_dlog_;
            kc.dlog_fill = d_dlog_fill_;
            kc.dlog_next = d_dlog_next_;
            kc.dlog_segs = (uint32_t)(dlog_records_ / kLogSeg);
            hipLaunchKernelGGL(dev::zero_u64, dim3(1), dim3(64), 0, s, d_cursor_, 2);
            hipLaunchKernelGGL(dev::zero_u64, dim3(1), dim3(64), 0, s, d_dlog_next_, 1);
            HIP_OK(hipGetLastError());
            launch_log(kc, cblocks, cblock, lds_bytes(kc.tile_n), s);
            HIP_OK(hipGetLastError());
            enqueue_log_passes(s, (uint64_t)((double)n * rpp));
            done += n;
        }
        HIP_OK(hipEventRecord((hipEvent_t)ev1_, s));
        HIP_OK(hipStreamSynchronize(s));
        float ms = 0.f;
        HIP_OK(hipEventElapsedTime(&ms, (hipEvent_t)ev0_, (hipEvent_t)ev1_));
        st.kernel_ms = ms;
        return st;
    }
    hipLaunchKernelGGL(dev::zero_u64, dim3(1), dim3(64), 0, s, d_cursor_, 2);
    HIP_OK(hipGetLastError());
    HIP_OK(hipEventRecord((hipEvent_t)ev0_, s));
    launch(ka, rblocks, rblock, lds, s);
    HIP_OK(hipGetLastError());
    HIP_OK(hipEventRecord((hipEvent_t)ev1_, s));
    HIP_OK(hipStreamSynchronize(s));
    float ms = 0.f;
    HIP_OK(hipEventElapsedTime(&ms, (hipEvent_t)ev0_, (hipEvent_t)ev1_));
    st.kernel_ms = ms;
    return st;
}

void Session::set_voxel_props(const float* mua, const float* mus, bool on_device) {
    HIP_OK(hipSetDevice(lc_.device));
    hipStream_t s = (hipStream_t)stream_;
    const uint64_t n = (uint64_t)prob_.nx * prob_.ny * prob_.nz;
    if (!mua || !mus) {
        if (d_vprops_) HIP_OK(hipFree(d_vprops_));
        d_vprops_ = nullptr;
        return;
    }
    if (!d_vprops_) d_vprops_ = dalloc<float2>(n);
    const float *dm = mua, *ds = mus;
    float* staging = nullptr;
    if (!on_device) {
        staging = dalloc<float>(2 * n);
        HIP_OK(hipMemcpyAsync(staging, mua, n * 4, hipMemcpyHostToDevice, s));
        HIP_OK(hipMemcpyAsync(staging + n, mus, n * 4, hipMemcpyHostToDevice, s));
        dm = staging;
        ds = staging + n;
    }
    hipLaunchKernelGGL(dev::vprops_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, dm, ds, prob_.unitinmm,
                       (float2*)d_vprops_, n);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(s));
    if (staging) HIP_OK(hipFree(staging));
    tile_chosen_ = false;
    for (int& v : tile_) v = 0;
    dep_cache_ = true;
    dlog_auto_ = false;
    dlog_rpp_ = 0.0;
    vote_auto_ = 0;
}

void Session::set_sources(const float* table, int nsrc, uint64_t per_source) {
    HIP_OK(hipSetDevice(lc_.device));
    if (d_msrc_) HIP_OK(hipFree(d_msrc_));
    d_msrc_ = nullptr;
    msrc_n_ = 0;
    msrc_count_ = 0;
    if (!table || nsrc <= 0 || per_source == 0) return;
    if ((uint64_t)nsrc * per_source > 0xFFFFFFFFull)
        throw std::runtime_error("several sources in one run: at most 2^32 photons in total");
    d_msrc_ = dalloc<float>((size_t)nsrc * 6);
    HIP_OK(hipMemcpy(d_msrc_, table, (size_t)nsrc * 6 * 4, hipMemcpyHostToDevice));
    msrc_n_ = per_source;
    msrc_count_ = (uint64_t)nsrc;
    tile_chosen_ = false;
    for (int& v : tile_) v = 0;
    dep_cache_ = true;
    dlog_auto_ = false;
    dlog_rpp_ = 0.0;
    vote_auto_ = 0;
}

void Session::set_source(const float pos[3], const float dir[3]) {
    for (int i = 0; i < 3; i++) {
        prob_.srcpos[i] = pos[i];
        prob_.srcdir[i] = dir[i];
    }
    tile_chosen_ = false;
    for (int& v : tile_) v = 0;
    dep_cache_ = true;
    dlog_auto_ = false;
    dlog_rpp_ = 0.0;
    vote_auto_ = 0;
}

RunStats Session::replay(const uint32_t* ids, const int32_t* slices, const float* coef, uint64_t n, float* gmua,
                         float* gmus, uint64_t stride, float* final_w, const int32_t* src_index,
                         const float* sources) {
    HIP_OK(hipSetDevice(lc_.device));
    hipStream_t s = (hipStream_t)stream_;
    RunStats st;
    st.photons = n;
    if (n == 0) return st;
    dev::KernelArgs ka;
    build_args(0, n, &ka);
    ka.trace = nullptr;
    ka.detnum = 0; ka.ppathlen = 0; ka.issavevol = 0;
    ka.tile_n = 0; ka.tsx = ka.tsy = ka.tsz = ka.tsg = 0;
    ka.rids = ids; ka.rslice = slices; ka.rcoef = coef;
    ka.gmua = gmua; ka.gmus = gmus; ka.gstride = stride; ka.rw = final_w;
    ka.rsrc = sources ? src_index : nullptr;
    ka.rsrcs = sources;
    ka.msrc_n = 0;
    ka.vote = 0;
    hipLaunchKernelGGL(dev::zero_u64, dim3(1), dim3(64), 0, s, d_cursor_, 2);
    HIP_OK(hipGetLastError());
    HIP_OK(hipEventRecord((hipEvent_t)ev0_, s));
    pick_replay(prob_.isreflect != 0, label_bytes_ == 1, label_mode())(ka, blocks_, block_, lds_bytes(0, false), s);
    HIP_OK(hipGetLastError());
    HIP_OK(hipEventRecord((hipEvent_t)ev1_, s));
    HIP_OK(hipStreamSynchronize(s));
    float ms = 0.f;
    HIP_OK(hipEventElapsedTime(&ms, (hipEvent_t)ev0_, (hipEvent_t)ev1_));
    st.kernel_ms = ms;
    return st;
}

void Session::fetch_fluence(float* out) const {
    HIP_OK(hipSetDevice(lc_.device));
    if (fluence_len_) HIP_OK(hipMemcpy(out, d_fluence_, fluence_len_ * 4, hipMemcpyDeviceToHost));
}

uint32_t Session::detected() const {
    HIP_OK(hipSetDevice(lc_.device));
    uint32_t n = 0;
    HIP_OK(hipMemcpy(&n, d_counters_, 4, hipMemcpyDeviceToHost));
    return n;
}

uint32_t Session::saved() const { return std::min(detected(), detcap_); }

void Session::fetch_detp(float* rows, uint32_t* ids, float* weights) const {
    HIP_OK(hipSetDevice(lc_.device));
    uint32_t n = saved();
    int cols = prob_.detp_columns();
    if (rows && n && cols) HIP_OK(hipMemcpy(rows, d_detp_, (size_t)n * cols * 4, hipMemcpyDeviceToHost));
    if (ids && n) HIP_OK(hipMemcpy(ids, d_detid_, (size_t)n * 4, hipMemcpyDeviceToHost));
    if (weights && n) HIP_OK(hipMemcpy(weights, d_detw_, (size_t)n * 4, hipMemcpyDeviceToHost));
}

void Session::energies(double out[3]) const {
    HIP_OK(hipSetDevice(lc_.device));
    HIP_OK(hipMemcpy(out, d_energy_, 24, hipMemcpyDeviceToHost));
}

void Session::fetch_trace(TraceRec* out, uint64_t n) const {
    HIP_OK(hipSetDevice(lc_.device));
    if (!d_trace_) throw std::runtime_error("trace was not enabled for this session");
    n = std::min(n, last_count_);
    HIP_OK(hipMemcpy(out, d_trace_, n * sizeof(TraceRec), hipMemcpyDeviceToHost));
}

uint64_t device_arith_check(int which, uint64_t* tested, int device) {
    HIP_OK(hipSetDevice(device));
    const uint64_t n = which == 0 ? (0x7F800000ull - 0x0F800000ull)
                     : which == 1 ? 2ull * (0x5D800001ull - 0x21800000ull)
                     : which == 3 ? 2ull * (0x7E800000ull - 0x00800000ull)
                     : which == 4 ? 2ull * 8ull * (1ull << 23) : (1ull << 32);
    unsigned long long* bad = dalloc<unsigned long long>(1);
    HIP_OK(hipMemset(bad, 0, 8));
    hipLaunchKernelGGL(dev::arith_check_kernel, dim3(8192), dim3(256), 0, 0, which, n, bad);
    HIP_OK(hipGetLastError());
    unsigned long long h = 0;
    HIP_OK(hipMemcpy(&h, bad, 8, hipMemcpyDeviceToHost));
    (void)hipFree(bad);
    if (tested) *tested = n;
    return h;
}

TallyBoxSearch device_choose_box(const float* counts, int nx, int ny, int nz, int nt, uint64_t budget, int device) {
    HIP_OK(hipSetDevice(device));
    const uint64_t n = (uint64_t)nx * ny * nz * (uint64_t)nt;
    float* d_cnt = dalloc<float>(n);
    unsigned long long* d_s = dalloc<unsigned long long>(8);
    HIP_OK(hipMemcpy(d_cnt, counts, n * 4, hipMemcpyHostToDevice));
    enqueue_box_search(d_cnt, n, nt, nx, ny, nz, budget, d_s, nullptr);
    unsigned long long h[8];
    HIP_OK(hipMemcpy(h, d_s, sizeof h, hipMemcpyDeviceToHost));
    (void)hipFree(d_cnt);
    (void)hipFree(d_s);
    TallyBoxSearch r;
    memcpy(&r.total, &h[2], 8);
    memcpy(&r.in_box, &h[3], 8);
    memcpy(r.box, &h[4], sizeof r.box);
    return r;
}

// ---------------- one photon range on caller-owned buffers, with no allocation ----------------
//
// umcx_run_device_noalloc runs the kernel of umcx_session_run on buffers the caller owns. What a
// session allocates beside its inputs and outputs -- the work cursor and the packed media ids
// staged in LDS -- lives here in one file-scope device array, `scratch`, that every call on a
// device shares (calls that overlap on one device race on it). Its address on each device comes
// from hipGetSymbolAddress, taken once per device with that device current; host code uses only
// that address, never the array's name.
constexpr size_t kNoAllocCtlBytes = 256;                                  // work cursor (2 x u64), padding
constexpr size_t kNoAllocScratchBytes = kNoAllocCtlBytes + kLabelCacheBytes;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
struct NoAllocDevice {
    bool ready = false;
    unsigned char* ptr = nullptr;   // this device's copy of `scratch`
    int cus = 0;
    size_t lds_wg_max = 0;
};
constexpr int kNoAllocMaxDevices = 64;
NoAllocDevice g_noalloc[kNoAllocMaxDevices];
std::mutex g_noalloc_mu;

const NoAllocDevice& noalloc_device(int dev) {
    if (dev < 0 || dev >= kNoAllocMaxDevices) throw std::runtime_error("umcx_run_device_noalloc: device ordinal out of range");
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    NoAllocDevice& d = g_noalloc[dev];
    if (!d.ready) {
        void* ptr = nullptr;
        HIP_OK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        int cus = 0, lds = 0;
        HIP_OK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        HIP_OK(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        d.ptr = (unsigned char*)ptr;
        d.cus = std::max(1, cus);
        d.lds_wg_max = (lds > 0 ? (size_t)lds : 64 * 1024) - kStaticLds;
        d.ready = true;
    }
    return d;
}
}  // namespace

void run_device_noalloc(const umcx_device_problem* q, uint64_t first, uint64_t count, void* stream) {
    const char* fn = "umcx_run_device_noalloc";
    if (!q) throw std::runtime_error(std::string(fn) + ": null problem");
    const uint32_t nx = q->dims[0], ny = q->dims[1], nz = q->dims[2];
    const int32_t maxgate = (int32_t)q->dims[3];
    if (!nx || !ny || !nz || maxgate < 1) throw std::runtime_error(std::string(fn) + ": empty grid or no time gate");
    const uint64_t nvox = (uint64_t)nx * ny * nz;
    if (nvox >= 0x7FFFFFFFull) throw std::runtime_error(std::string(fn) + ": more than 2^31 - 1 voxels");
    if (q->mediumnum < 1 || q->mediumnum > 16)
        throw std::runtime_error(std::string(fn) + ": mediumnum must be 1..16 (media ids are staged at 2 or 4 bits)");
    if (q->srctype < 0 || q->srctype >= (int32_t)kUnknownSource) throw std::runtime_error(std::string(fn) + ": unknown source type");
    if (!q->labels || !q->media || !q->energies) throw std::runtime_error(std::string(fn) + ": null labels, media or energies");
    if (first > ~0ull - count) throw std::runtime_error(std::string(fn) + ": photon range overflows");
    const uint32_t bits = q->mediumnum <= 4 ? 2u : 4u;
    const uint64_t words = (nvox * bits + 31) / 32;
    if (words * 4 > kLabelCacheBytes)
        throw std::runtime_error(std::string(fn) + ": the packed label volume needs " + std::to_string(words * 4) +
                                 " bytes of scratch and at most " + std::to_string(kLabelCacheBytes) +
                                 " are available (" + std::to_string(nvox) + " voxels at " + std::to_string(bits) +
                                 " bits); run it through a session (umcx_session_run)");
    const bool savevol = q->fluence != nullptr;
    const uint64_t cells = nvox * (uint64_t)maxgate;
    if (savevol && q->deposit_cache && cells >= 0xFFFFFFFFull)
        throw std::runtime_error(std::string(fn) + ": the deposit cache addresses fewer than 2^32 - 1 fluence cells");
    const bool savedet = q->issavedet != 0;
    const int32_t detnum = savedet ? q->detnum : 0;
    const int32_t ppathlen = (savedet && (q->savedetflag & kPPath)) ? q->mediumnum : 0;
    const int32_t detlen = (q->savedetflag & kDetID ? 1 : 0) + ((q->savedetflag & kPPath) ? q->mediumnum : 0) +
                           (((q->savedetflag & kExitPos) ? 1 : 0) + ((q->savedetflag & kExitDir) ? 1 : 0)) * 3;
    if (detnum < 0) throw std::runtime_error(std::string(fn) + ": negative detector count");
    if (savedet && (!q->ndet || !q->detid || !q->detw || (detlen > 0 && !q->detp) || (detnum > 0 && !q->detpos)))
        throw std::runtime_error(std::string(fn) + ": detectors need detpos, detp, detid, detw and ndet");
    if (count == 0) return;

    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    const NoAllocDevice& d = noalloc_device(dev);
    hipStream_t s = (hipStream_t)stream;
    unsigned long long* cursor = (unsigned long long*)d.ptr;
    uint32_t* plabels = (uint32_t*)(d.ptr + kNoAllocCtlBytes);

    // launch shape of the sessions with staged labels: one 1024-lane workgroup per CU, half-size
    // workgroups for runs under kShortRunPhotonsPerLane photons per lane
    int block = 1024;
    if (count < kShortRunPhotonsPerLane * 1024ull * (uint64_t)d.cus) block = 512;
    auto lds_of = [&](int b) {
        return (size_t)32 * 16 + (size_t)q->mediumnum * 16 + (size_t)words * 4 +
               (ppathlen > kRegPaths ? (size_t)ppathlen * b * 4 : 0);
    };
    while (lds_of(block) > d.lds_wg_max && block > 64) block /= 2;
    if (lds_of(block) > d.lds_wg_max) throw std::runtime_error(std::string(fn) + ": media and labels exceed the LDS");
    const int blocks = d.cus;

    dev::KernelArgs ka;
    memset(&ka, 0, sizeof ka);
    ka.nx = nx; ka.ny = ny; ka.nz = nz; ka.nyz = ny * nz; ka.nxyz = nvox;
    ka.maxgate = maxgate; ka.outputtype = q->outputtype; ka.savedetflag = q->savedetflag;
    ka.mediumnum = q->mediumnum; ka.detnum = detnum; ka.srctype = q->srctype;
    ka.issavevol = savevol ? 1 : 0; ka.ppathlen = ppathlen; ka.detlen = detlen;
    ka.tstart = q->tstart; ka.tend = q->tend; ka.rtstep = q->rtstep;
    ka.srcpos = make_float4(q->srcpos[0], q->srcpos[1], q->srcpos[2], q->srcpos[3]);
    ka.srcdir = make_float4(q->srcdir[0], q->srcdir[1], q->srcdir[2], q->srcdir[3]);
    ka.srcp1 = make_float4(q->srcparam1[0], q->srcparam1[1], q->srcparam1[2], q->srcparam1[3]);
    ka.srcp2 = make_float4(q->srcparam2[0], q->srcparam2[1], q->srcparam2[2], q->srcparam2[3]);
    for (int i = 0; i < 4; i++) ka.seeds[i] = q->seeds[i];
    ka.first = first; ka.count = count;
    ka.labels = q->labels;
    ka.media = (const float4*)q->media;
    ka.detpos = (const float4*)q->detpos;
    ka.fluence = q->fluence;
    ka.detp = q->detp; ka.detid = q->detid; ka.detw = q->detw; ka.ndet = q->ndet;
    ka.maxdet = savedet ? q->maxdetphotons : 0u;
    ka.energy = q->energies;
    ka.cursor = cursor;
    ka.dep_cache = (savevol && q->deposit_cache) ? 1 : 0;
    ka.unit = q->unitinmm;
    ka.plabels = plabels;
    ka.plabel_words = (uint32_t)words;
    ka.plabel_bits = bits;
    ka.plabel_lg = bits == 2 ? 4u : 3u;
    ka.plabel_id0 = 0;
    ka.plabel_step = 1;
    ka.per_wg = (count - count / 8) / (uint64_t)blocks;

    // the cursor is cleared by a fill kernel (not a memset node, so the sequence can be captured
    // in a graph) and the packed labels are rewritten in full by every call
    hipLaunchKernelGGL(dev::zero_u64, dim3(1), dim3(64), 0, s, cursor, 2);
    HIP_OK(hipGetLastError());
    hipLaunchKernelGGL(dev::pack_labels_kernel, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, s, q->labels, nvox,
                       bits, (uint32_t)words, plabels);
    HIP_OK(hipGetLastError());
    pick(q->isreflect != 0, savedet, true, dev::kLabLds)(ka, blocks, block, lds_of(block), s);
    HIP_OK(hipGetLastError());
}

void device_libm_eval(int fn, const float* in, float* out, uint64_t n, int device) {
    HIP_OK(hipSetDevice(device));
    float *din = dalloc<float>(n), *dout = dalloc<float>(n);
    HIP_OK(hipMemcpy(din, in, n * 4, hipMemcpyHostToDevice));
    unsigned blocks = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(dev::libm_eval_kernel, dim3(blocks), dim3(256), 0, 0, fn, din, dout, n);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpy(out, dout, n * 4, hipMemcpyDeviceToHost));
    (void)hipFree(din);
    (void)hipFree(dout);
}

}  // namespace umcx
