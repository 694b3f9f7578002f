/*
 * umcx_oracle.h -- CPU restatement of umcx's photon loop (TEST INFRASTRUCTURE ONLY).
 *
 * This is the checker, not the product. Only tests/, the smoke test and the
 * cpu_baseline leg of bench.py may load it. The product path (umcx_amd/ + libumcx.so)
 * never links, loads or calls anything here.
 *
 * It follows /src/umcx.cpp of fangq/umcx and calls the host C library (glibc libm and
 * the reference's seeding scheme) exactly where the reference does, so that on the same
 * host it reproduces the reference CPU build's per-photon trajectories bit for bit.
 * Every function cites the reference lines it restates.
 */
#ifndef UMCX_ORACLE_H
#define UMCX_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Flattened simulation constants, restating MCX_param (src/umcx.cpp:95-99) plus the
 * source and seeding inputs that MCX_kernel reads from the JSON (src/umcx.cpp:725-729). */
typedef struct oracle_problem {
    uint32_t nx, ny, nz;          /* label volume dims; memory order [x][y][z], z fastest (:128-130) */
    int32_t maxgate;              /* number of time gates (:789) */
    float tstart, tend, rtstep, unitinmm;
    int32_t isreflect, issavevol, issavedet, savedetflag;
    int32_t mediumnum, outputtype, detnum, srctype;
    uint32_t maxdetphotons;
    float srcpos[4], srcdir[4], srcparam1[4], srcparam2[4];
    uint32_t seeds[4];            /* the four glibc rand() words (:725-727) */
    const int32_t* labels;        /* nx*ny*nz labels, sign bit = detector mask (:35, :605-632) */
    const float* media;           /* mediumnum x {mua, mus, g, n}, already in per-voxel units (:803-806) */
    const float* detpos;          /* detnum x {x, y, z, R} (:808-810) */
} oracle_problem;

/* Final state of one photon, for per-photon parity and the work model. */
typedef struct oracle_trace {
    float pos[4];                 /* exit position, pos.w = final weight */
    float vec[4];                 /* exit direction, vec.w = scatter count (:379) */
    float len[4];                 /* {remaining scattering length, time of flight, path since save, w at last save} */
    int32_t mediaid, lastvoxelidx;
    float w0;                     /* launch weight (:762) */
    uint32_t nstep, ncross, nscat, nrand, nfresnel;
} oracle_trace;

/* glibc srand()/rand() TYPE_3 restatement: writes the four words MCX_kernel draws (:725-727). */
void oracle_seed_words(int32_t seed, uint32_t out[4]);

/* Ground-truth reference of libm's logf/sinf/cosf/acosf/expf, i.e. the host C library. */
float oracle_host_logf(float x);
float oracle_host_sinf(float x);
float oracle_host_cosf(float x);
float oracle_host_acosf(float x);
float oracle_host_expf(float x);
/* fn: 0 logf, 1 sinf, 2 cosf, 3 acosf, 4 expf of the host C library over an array. */
void oracle_host_libm_array(int fn, const float* in, float* out, uint64_t n);

/*
 * Run photons [first, first+count) of the MCX_kernel loop (:752-775).
 *   fluence   : nx*ny*nz*maxgate floats, accumulated into (not cleared), may be NULL if !issavevol
 *   detp      : maxdetphotons x detphotondatalen floats (row layout of :157, :164-176), may be NULL
 *   detp_id   : photon index per saved row (extension; the reference saves no id), may be NULL
 *   detp_w    : final weight per saved row (extension), may be NULL
 *   ndetected : incremented by every detection, including dropped ones (:166-169)
 *   energies  : [0] += escaped weight, [1] += launched weight (:765-770), [2] += sum of deposits
 *   traces    : count entries of per-photon final state, may be NULL
 *   nthreads  : OpenMP threads (<=0: library default)
 * Returns 0, or -1 on invalid input.
 */
int oracle_run(const oracle_problem* p, uint64_t first, uint64_t count, float* fluence,
               float* detp, uint32_t* detp_id, float* detp_w, uint32_t* ndetected,
               double energies[3], oracle_trace* traces, int nthreads);

/* Replay of detected photons for the differentiable form: per-voxel perturbation-MC terms of
 * record k, in voxel v of the path, added into slice slices[k] (stride floats apart):
 *   gmua += coef[k] * (-L_v)     gmus += coef[k] * (n_v / mus_v - L_v)
 * L_v: path length in voxel v (mm), n_v: scatterings in v, mus_v in 1/mm. Photon ids[k] is
 * launched from source src_index[k] of sources ({pos xyz, dir xyz}; NULL: the problem's).
 * amua (sums of |mua term|), amus (sums of |coef| (n_v / mus_v + L_v)), final_w, mpath
 * (per-medium float segment sums, mediumnum per record) and vprops
 * (per-voxel {mua, mus} per voxel length) may be NULL. Returns 0, or -1 on invalid input. */
int oracle_replay(const oracle_problem* p, const uint32_t* ids, const int32_t* src_index, const float* sources,
                  const int32_t* slices, const double* coef, uint64_t n, const float* vprops, double* gmua,
                  double* gmus, double* amua, double* amus, uint64_t stride, float* final_w, float* mpath,
                  int nthreads);

/* Select a planted mistake (0 = none); used only by tests that prove the checks have teeth.
 * 4 credits each replay segment to the voxel the photon enters instead of the one it leaves. */
void oracle_set_mutant(int m);

/* Length of a detected-photon row for this problem (:157). */
int oracle_detp_len(const oracle_problem* p);

#ifdef __cplusplus
}
#endif
#endif
