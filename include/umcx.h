/*
 * umcx.h -- C ABI of the MI355X umcx photon-transport library (libumcx.so).
 *
 * Drop-in part: the two C exports of fangq/umcx (src/umcx.cpp:850-856), same names, arguments
 * and outputs (files <ID>.bnii / <ID>_detp.jdb in the working directory plus one summary line
 * on stdout). C++ callers also get MCX_run_simulation with the reference's signature
 * (src/umcx.cpp:783). Unlike the reference, no C++ exception crosses this boundary: every entry
 * point returns 0 on success and non-zero on failure, with the message in umcx_last_error().
 *
 * Session part: the photon loop MCX_kernel (src/umcx.cpp:722-781) as a reusable GPU session,
 * so a caller can prepare a problem once, run photon ranges on the GPU, and read tallies,
 * detected photons and energies back without going through files.
 *
 * Threads: the drop-in exports and separate sessions or batches may be used from several threads
 * at once (each session and batch owns its device buffers and stream; error messages are
 * per thread). One session or batch handle must not be used by two threads at the same time.
 */
#ifndef UMCX_H
#define UMCX_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- drop-in exports ---------------------------------------------------------------- */

/* Replaces MCX_run_cmd (src/umcx.cpp:850): argv/argn as for the umcx command line. */
int MCX_run_cmd(char* argv[], int argn);
/* Replaces MCX_run_json (src/umcx.cpp:853): one JSON document merged into an empty config. */
int MCX_run_json(char* jsoninput);

/* Message of the last failure on this thread ("" when none). */
const char* umcx_last_error(void);
const char* umcx_version(void);
int umcx_device_count(void);

/* ---- GPU session ---------------------------------------------------------------------- */

typedef struct umcx_session umcx_session;

/* Launch options. A NULL pointer takes the tuned defaults for every field. In a caller-filled
 * struct, block, blocks, label_cache, deposit_cache and deposit_log take their tuned defaults at 0, but
 * tally_mode and vote must be -1 for theirs (0 selects global atomics only and immediate
 * scattering). No counterpart in the reference, whose ThreadNum/BlockSize session keys
 * (src/umcx.cpp:739-741) are honoured when block and blocks are zero. */
typedef struct umcx_launch {
    int32_t device;      /* HIP device ordinal */
    int32_t block;       /* threads per workgroup */
    int32_t blocks;      /* persistent workgroups */
    int32_t tally_mode;  /* -1 tuned, 0 global atomics, 1 LDS tile */
    int32_t trace;       /* 1: keep every photon's final state (small runs) */
    int32_t vote;        /* scatter-round threshold in lanes; 0 scatters as soon as a path ends, -1 tuned */
    int32_t label_cache; /* 0: media ids packed into LDS when the volume fits (<= 96 KiB at 2 or 4 bits
                            per voxel; 1 bit after the pilot for at most two media ids, see
                            umcx_info.lds_label_bits), -1: labels always read from HBM */
    int32_t deposit_cache; /* 0: per-lane cache of the last 4 global fluence deposits when the LDS tile
                              takes under half of them (pilot count), 1: always, -1: never */
    int32_t deposit_log;   /* 0: global fluence deposits appended to a device log and summed into the
                              grid by separate passes when the pilot counts at least 200 of them per
                              photon (float atomics would saturate), 1: always, -1: never (float
                              atomics); grids of 2^27 cells or more always use atomics */
    int32_t deposit_log_segments; /* log capacity in segments of 4096 records, 64 KiB of device memory
                                     each with the grouped copy (0: sized per run from the pilot's
                                     count, at most a quarter of the free device memory); a run that
                                     outgrows the log runs in chunks, and deposits past a full log
                                     go to float atomics */
} umcx_launch;

/* Constants of a prepared problem: MCX_param (src/umcx.cpp:95-99) plus launch facts. */
typedef struct umcx_info {
    uint32_t dims[4];    /* nx, ny, nz, maxgate */
    uint64_t nphoton;
    int32_t mediumnum, detnum, detp_columns, ppath_columns;
    int32_t outputtype, isnormalized, issavevol, issavedet, isreflect, srctype;
    float tstart, tend, rtstep, unitinmm;
    uint32_t seeds[4];
    int32_t block, blocks, tally_mode, label_bytes;
    int32_t lds_label_bits;  /* bits per voxel of the LDS-staged labels (0: labels read from HBM);
                                a volume of at most two media ids may switch from 2 or 4 to 1
                                after the pilot, for a larger tally tile */
    int32_t last_block, last_blocks;  /* launch shape of the last run (0 before any): with no explicit
                                         shape, runs under 12 photons per lane of the full shape use
                                         half the lanes */
    int32_t deposit_cache;            /* 1 when runs use the per-lane deposit cache */
    double deposit_cache_hits;        /* share of the pilot's global deposits that hit that cache */
    int32_t deposit_log;              /* 1 when runs append global deposits to the deposit log */
    double global_deposits_per_photon; /* the pilot's estimate for the runs (0 before a pilot) */
    int32_t vote;                     /* scatter-round lane threshold the runs use (0: none); with
                                         umcx_launch.vote -1 the pilot picks 16 when it counts fewer
                                         than 0.3 scatters per voxel crossing, else 0 */
} umcx_info;

/* Build a session from a complete umcx JSON config (the MCX_userio input, src/umcx.cpp:451-632)
 * or from command-line arguments (argn < 0: argv[0] is a BJData config buffer of -argn bytes,
 * :475-478). Uploads the domain to the GPU. */
int umcx_session_from_json(const char* json, const umcx_launch* launch, umcx_session** out);
int umcx_session_from_argv(char** argv, int argn, const umcx_launch* launch, umcx_session** out);
int umcx_session_info(const umcx_session* s, umcx_info* info);
/* Zero the fluence grid, detected photons and energy tallies. */
int umcx_session_reset(umcx_session* s);
/* Run photons [first, first+count) of the MCX_kernel loop (src/umcx.cpp:752-775) and wait;
 * tallies accumulate. kernel_ms (may be NULL) receives the HIP-event time of the kernel. */
int umcx_session_run(umcx_session* s, uint64_t first, uint64_t count, double* kernel_ms);
/* Raw (unnormalised) fluence tallies, memory order [t][x][y][z]; n = dims product. */
int umcx_session_fluence(const umcx_session* s, float* out, uint64_t n);
/* Device pointer of the fluence grid (stays owned by the session). */
int umcx_session_fluence_device(const umcx_session* s, float** dptr);
/* All detections (incl. ones beyond MaxDetPhoton) and the number of rows kept (:182-184). */
int umcx_session_detected(const umcx_session* s, uint32_t* detected, uint32_t* saved);
/* Detected-photon rows [saved x detp_columns] in the layout of src/umcx.cpp:164-176, plus the
 * photon index and final weight of each row (both extensions). Any pointer may be NULL. */
int umcx_session_detp(const umcx_session* s, float* rows, uint32_t* ids, float* weights);
/* {escaped, launched, deposited} weight sums in fp64 (src/umcx.cpp:765-770). */
int umcx_session_energies(const umcx_session* s, double out[3]);
/* Per-photon final state of the last run (64 bytes each), when launched with trace = 1. */
int umcx_session_trace(const umcx_session* s, void* out, uint64_t n);
/* ---- differentiation (no counterpart in the reference) ----------------------------------- */

/* Per-voxel absorption and scattering coefficients in 1/mm (nx*ny*nz each, [x][y][z] order)
 * overriding the labels' mua/mus; g and n stay per label. Device pointers when on_device = 1,
 * host pointers otherwise; NULL restores the labels' values. With values equal to the labels'
 * media the simulation is bitwise unchanged. */
int umcx_session_set_voxel_props(umcx_session* s, const float* mua, const float* mus, int32_t on_device);
/* Move the source (position and direction, grid units) for the following runs. */
int umcx_session_set_source(umcx_session* s, const float pos[3], const float dir[3]);
/* Device pointers of the detected-photon arrays (photon index, final weight, rows). */
int umcx_session_detp_device(const umcx_session* s, uint32_t** ids, float** weights, float** rows);
/* Replay: re-trace photons ids[0..n) (device arrays) with the session's current source and
 * properties, adding for record k, in every voxel v it crosses,
 *   grad_mua[slices[k]*slice_stride + v] += coef[k] * (-L_v)
 *   grad_mus[slices[k]*slice_stride + v] += coef[k] * (n_v / mus_v - L_v)
 * (L_v path length in mm, n_v scatterings, mus_v in 1/mm): with coef = w / I these are
 * d log I / d mua_v and d log I / d mus_v of a detected intensity I = sum of final weights w
 * (perturbation Monte Carlo). final_w (device, may be NULL) receives each replayed photon's
 * final weight. All pointers are device pointers. */
int umcx_session_replay(umcx_session* s, const uint32_t* ids, const int32_t* slices, const float* coef, uint64_t n,
                        float* grad_mua, float* grad_mus, uint64_t slice_stride, float* final_w, double* kernel_ms);

/* Several sources in one run: sources = {pos xyz, dir xyz} per source (host), photon slot p of
 * umcx_session_run is photon p mod photons_per_source of source p / photons_per_source, seeded
 * and traced exactly as in that source's own run; detection records carry the slot. While a
 * table is set, umcx_session_run rejects (before any launch) a range [first, first+count) that
 * ends past nsrc * photons_per_source; nsrc * photons_per_source is at most 2^32. NULL or
 * nsrc 0 returns to the single source. No counterpart in the reference. */
int umcx_session_set_sources(umcx_session* s, const float* sources, int32_t nsrc, uint64_t photons_per_source);
/* umcx_session_replay for the records of several sources in one launch: record k is launched
 * from source j = src_index[k], whose position and direction are sources[6j..6j+5] = {pos xyz,
 * dir xyz} (all device pointers); the source type and its parameters are the session's. No
 * counterpart in the reference. */
int umcx_session_replay_sources(umcx_session* s, const uint32_t* ids, const int32_t* slices, const float* coef,
                                const int32_t* src_index, const float* sources, uint64_t n, float* grad_mua,
                                float* grad_mus, uint64_t slice_stride, float* final_w, double* kernel_ms);

/* The LDS tally tile in use, {x0, y0, z0, gate0, nx, ny, nz, ngates} (zeros when none).
 * No counterpart in the reference. */
int umcx_session_tally_box(const umcx_session* s, int32_t box[8]);
/* Output normaliser of src/umcx.cpp:820 for a given launched energy (NaN for a null session). */
float umcx_session_normalizer(const umcx_session* s, double energytotal);
void umcx_session_free(umcx_session* s);

/* ---- batched simulations (no counterpart in the reference) ------------------------------- */

/* Many independent simulations (each a complete umcx JSON config) run in one persistent launch
 * replayed from a hipGraph. Per-job results (fluence, energies, detected photons) equal separate
 * runs up to float summation order; detected rows equal them as a multiset. */
typedef struct umcx_batch umcx_batch;
int umcx_batch_create(const char* const* jsons, int32_t njobs, const umcx_launch* launch, umcx_batch** out);
/* Zero the tallies and run every job's photons; kernel_ms (may be NULL) gets the graph's time. */
int umcx_batch_run(umcx_batch* b, double* kernel_ms);
int umcx_batch_info(const umcx_batch* b, int32_t job, umcx_info* info);
int umcx_batch_fluence(const umcx_batch* b, int32_t job, float* out, uint64_t n);
int umcx_batch_energies(const umcx_batch* b, int32_t job, double out[3]);
/* Detected photons of one job (src/umcx.cpp:155-184): all detections and the rows kept (at most
 * the job's MaxDetPhoton), then the rows [saved x detp_columns] in the layout of :164-176 with
 * the photon index and final weight of each (any pointer may be NULL), as umcx_session_detected /
 * umcx_session_detp give them for a separate run. */
int umcx_batch_detected(const umcx_batch* b, int32_t job, uint32_t* detected, uint32_t* saved);
int umcx_batch_detp(const umcx_batch* b, int32_t job, float* rows, uint32_t* ids, float* weights);
/* One launch shape the autotuner timed at creation: workgroup size, persistent workgroups, LDS
 * tally cells per workgroup (0: global atomics only), the trial's kernel time over a 1/32 sample
 * of every job (0 when an explicit shape leaves a single candidate, which is not timed), the share
 * of pilot deposits inside the jobs' tiles, and whether it was chosen. */
typedef struct umcx_batch_trial {
    int32_t block, blocks, tile_cells, chosen;
    double ms, tile_share;
} umcx_batch_trial;
/* Writes at most cap trials to out (may be NULL) and their number to *n. */
int umcx_batch_trials(const umcx_batch* b, umcx_batch_trial* out, int32_t cap, int32_t* n);
/* LDS tally box of one job, {x0, y0, z0, gate0, nx, ny, nz, ngates} (zeros when none), and its
 * pilot deposits per photon. */
int umcx_batch_tally_box(const umcx_batch* b, int32_t job, int32_t box[8], double* deposits_per_photon);
void umcx_batch_free(umcx_batch* b);

/* ---- one photon range on caller-owned device buffers, with no allocation ------------------- */

/* A problem held entirely in caller-owned device memory. The constants are those of
 * umcx_session_info for the same config (media in per-voxel-length units, as MCX_param and the
 * media scaling of src/umcx.cpp:785-810 derive them); every pointer is a device pointer. */
typedef struct umcx_device_problem {
    uint32_t dims[4];          /* nx, ny, nz, maxgate */
    int32_t mediumnum;         /* media rows, label 0 included: 1..16 */
    int32_t detnum;            /* detector rows */
    int32_t isreflect, issavedet, savedetflag, outputtype, srctype;
    float tstart, tend, rtstep, unitinmm;
    float srcpos[4], srcdir[4], srcparam1[4], srcparam2[4];
    uint32_t seeds[4];         /* seed words (umcx_seed_words of RNGSeed) */
    uint32_t maxdetphotons;    /* rows the detp / detid / detw buffers hold */
    int32_t deposit_cache;     /* 1: per-lane cache of the last 4 global fluence deposits, 0: none */
    const uint8_t* labels;     /* nx*ny*nz, [x][y][z]: bits 0-6 medium id (< mediumnum), bit 7 detector mask */
    const float* media;        /* mediumnum x {mua, mus, g, n} */
    const float* detpos;       /* detnum x {x, y, z, r}, grid units */
    float* fluence;            /* maxgate*nx*ny*nz raw tallies [t][x][y][z], added to; NULL: no volume output */
    float* detp;               /* maxdetphotons x detp_columns rows (src/umcx.cpp:164-176) */
    uint32_t* detid;           /* photon index of each row */
    float* detw;               /* final weight of each row */
    uint32_t* ndet;            /* detections, incl. ones past maxdetphotons; added to */
    double* energies;          /* {escaped, launched, deposited}, added to */
} umcx_device_problem;

/* Photons [first, first+count) of umcx_session_run's kernel on the buffers of *p, enqueued on
 * `stream` (a hipStream_t of the current device; NULL: the null stream) without waiting and
 * without allocating anything: no hipMalloc, no hipFree, no memset node, so it may be captured in
 * a HIP graph. The caller zeroes the outputs before the first range. No counterpart in the
 * reference: it exists for callers that cannot allocate (graph capture, serving loops, harnesses
 * that forbid it). The work cursor and the LDS-staged media ids live in one static device buffer
 * of 98,560 bytes per GPU, shared by every call on that GPU: calls that overlap on one GPU race,
 * and volumes whose media ids packed at 2 bits (<= 4 media) or 4 bits (<= 16 media) per voxel
 * exceed 96 KiB are rejected (use a session). No LDS tally tile is used; results equal
 * umcx_session_run's up to float summation order. Errors are reported before anything is
 * enqueued. */
int umcx_run_device_noalloc(const umcx_device_problem* p, uint64_t first, uint64_t count, void* stream);

/* ---- host-only helpers (no GPU needed) ---------------------------------------------------- */

/* Label volume with detector bits, as MCX_userio::initdomain + maskdetectors build it
 * (src/umcx.cpp:537-632). *labels is allocated by the library; release with umcx_free. */
int umcx_domain_from_json(const char* json, uint32_t dims[3], int32_t** labels);
/* Expanded config after flags/presets/merges, as --dumpjson prints it (src/umcx.cpp:525-527);
 * with argn < 0, the JSON form of the BJData config buffer argv[0] (-argn bytes). */
int umcx_config_from_argv(char** argv, int argn, char** json_out);
/* The LDS tally box the sessions choose from a pilot's deposit counts (counts: ngates x nx x ny x
 * nz, memory order [t][x][y][z]): greedy growth from the densest cell, adding the face slab with
 * the most counts per added cell while the box holds at most `budget` cells; box = {x0, y0, z0,
 * gate0, nx, ny, nz, ngates}, all zero for an empty grid. No counterpart in the reference. */
int umcx_choose_tally_box(const float* counts, int32_t nx, int32_t ny, int32_t nz, int32_t ngates, uint64_t budget,
                          int32_t box[8]);
/* The same search as run on the GPU by the sessions (densest cell, total and greedy growth on the
 * device; the counts are copied there first); sums = {counts in the box, total} when not null.
 * No counterpart in the reference: it lets a caller check the device search against
 * umcx_choose_tally_box. */
int umcx_choose_tally_box_device(const float* counts, int32_t nx, int32_t ny, int32_t nz, int32_t ngates,
                                 uint64_t budget, int32_t box[8], double sums[2], int32_t device);
/* glibc srand(seed)+4*rand() words used to seed every photon (src/umcx.cpp:725-727). */
int umcx_seed_words(int32_t seed, uint32_t out[4]);
/* Evaluate the libm replicas (fn 0 logf, 1 sinf, 2 cosf, 3 acosf, 4 expf) on the host CPU. */
int umcx_libm_host(int32_t fn, const float* in, float* out, uint64_t n);
/* Same on the GPU (device ordinal), for bitwise checks of the device build. */
int umcx_libm_device(int32_t fn, const float* in, float* out, uint64_t n, int32_t device);
/* GPU self-check of the short correctly rounded sequences the photon loop uses away from special
 * operands (which 0: sqrt over every float in [2^-96, 2^128); 1: the direction reciprocal over
 * every float with 2^-60 <= |x| <= 2^60; 2: a/b on 2^32 hashed pairs of their domain; 3: the
 * reciprocal over every normal float with |x| < 2^126; 4: a/b for every b in [1, 2) against
 * eight edge mantissas of a) against the compiler's full sequences; *mismatches counts
 * differing results. No counterpart in the reference: it backs the bitwise claim of the device
 * arithmetic. */
int umcx_arith_check_device(int32_t which, uint64_t* tested, uint64_t* mismatches, int32_t device);
void umcx_free(void* p);

#ifdef __cplusplus
}
/* C++ linkage, as in the reference (src/umcx.cpp:783). argn < 0: argv[0] is a BJData buffer of
 * -argn bytes holding the whole config, as the reference's MATLAB binding passes it
 * (src/umcx.cpp:475-478, :860-865); nlhs/plhs (MATLAB outputs) are ignored and the output
 * files are written as for the other inputs. */
int MCX_run_simulation(char* argv[], int argn = 1, int nlhs = 0, void* plhs[] = nullptr);
#endif

#endif /* UMCX_H */
