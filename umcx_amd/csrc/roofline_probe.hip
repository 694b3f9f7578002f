// roofline_probe.hip -- measures the two gfx950 ceilings the photon kernels run against.
//
//   valu:    wave64 FP32 FMA issue rate (8 independent FMA chains per lane, 16 per unrolled step,
//            32 waves per SIMD over the run), in wave instructions per second;
//   atomics: device-scope float atomic adds (no return) to hashed addresses in a 40 MiB buffer
//            (the size of a time-gated fluence grid), in atomics per second, and the same into
//            one hot 256-float region (same-address serialisation at the memory side), and the
//            hashed rate over footprints from 64 KiB to 1 GiB, and with the nt / sc1 cache-policy bits.
//
//   --valu:  issue rates of the instruction classes the photon loop mixes (FP32 FMA and add,
//            32-bit integer, compare + select, register move, FP64 FMA, transcendental), each
//            from one probe kernel, in wave instructions per second.
//
// Prints JSON lines. Every index is taken modulo the buffer size; the launches are bounded.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>

#define CK(x)                                                                                  \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess) {                                                                \
            fprintf(stderr, "HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__);      \
            exit(1);                                                                           \
        }                                                                                      \
    } while (0)

__global__ void __launch_bounds__(256) fma_chains(float* out, int iters, float a, float b) {
    float x0 = threadIdx.x, x1 = x0 + 1, x2 = x0 + 2, x3 = x0 + 3, x4 = x0 + 4, x5 = x0 + 5, x6 = x0 + 6, x7 = x0 + 7;
    for (int i = 0; i < iters; i++) {
#pragma unroll
        for (int k = 0; k < 16; k++) {
            x0 = __builtin_fmaf(x0, a, b); x1 = __builtin_fmaf(x1, a, b);
            x2 = __builtin_fmaf(x2, a, b); x3 = __builtin_fmaf(x3, a, b);
            x4 = __builtin_fmaf(x4, a, b); x5 = __builtin_fmaf(x5, a, b);
            x6 = __builtin_fmaf(x6, a, b); x7 = __builtin_fmaf(x7, a, b);
        }
    }
    out[blockIdx.x * blockDim.x + threadIdx.x] = x0 + x1 + x2 + x3 + x4 + x5 + x6 + x7;
}

// Issue-rate probes, one instruction class each, for the instruction mix of the photon loop. Each
// lane runs 8 independent chains through 16 unrolled steps per iteration, so every step is 8
// instructions of the class (the loop counter runs on the scalar unit). The class of each loop
// body is checked in the gfx950 assembly (make -C umcx_amd/csrc probe_asm).
#define PROBE_BODY(OP)                                                                          \
    for (int i = 0; i < iters; i++) {                                                           \
        _Pragma("unroll") for (int k = 0; k < 16; k++) {                                        \
            OP(x0); OP(x1); OP(x2); OP(x3); OP(x4); OP(x5); OP(x6); OP(x7);                     \
        }                                                                                       \
    }

// v_add_f32
__global__ void __launch_bounds__(256) add_f32_chains(float* out, int iters, float a) {
    float x0 = threadIdx.x, x1 = x0 + 1, x2 = x0 + 2, x3 = x0 + 3, x4 = x0 + 4, x5 = x0 + 5, x6 = x0 + 6, x7 = x0 + 7;
#define OP_ADD(x) x = x + a
    PROBE_BODY(OP_ADD)
    out[blockIdx.x * blockDim.x + threadIdx.x] = x0 + x1 + x2 + x3 + x4 + x5 + x6 + x7;
}
// v_add_u32 and v_xor_b32, alternating
__global__ void __launch_bounds__(256) int32_chains(uint32_t* out, int iters, uint32_t a, uint32_t b) {
    uint32_t x0 = threadIdx.x, x1 = x0 * 3, x2 = x0 * 5, x3 = x0 * 7, x4 = x0 * 11, x5 = x0 * 13, x6 = x0 * 17, x7 = x0 * 19;
#define OP_INT(x) x = (x + a) ^ b
    PROBE_BODY(OP_INT)
    out[blockIdx.x * blockDim.x + threadIdx.x] = x0 ^ x1 ^ x2 ^ x3 ^ x4 ^ x5 ^ x6 ^ x7;
}
// v_cmp_gt_f32 + v_cndmask_b32: a compare into a lane mask and a select on it
__global__ void __launch_bounds__(256) select_chains(float* out, int iters, float a, float b) {
    float x0 = threadIdx.x, x1 = x0 + 1, x2 = x0 + 2, x3 = x0 + 3, x4 = x0 + 4, x5 = x0 + 5, x6 = x0 + 6, x7 = x0 + 7;
#define OP_SEL(x) x = (x > a) ? b : x
    PROBE_BODY(OP_SEL)
    out[blockIdx.x * blockDim.x + threadIdx.x] = x0 + x1 + x2 + x3 + x4 + x5 + x6 + x7;
}
// v_mov_b32 (register copies, as value selects and the deposit cache's shifts produce)
__global__ void __launch_bounds__(256) mov_chains(float* out, int iters) {
    float x0 = threadIdx.x, x1 = x0 + 1, x2 = x0 + 2, x3 = x0 + 3, x4 = x0 + 4, x5 = x0 + 5, x6 = x0 + 6, x7 = x0 + 7;
#define OP_MOV(x) asm volatile("v_mov_b32 %0, %0" : "+v"(x))
    PROBE_BODY(OP_MOV)
    out[blockIdx.x * blockDim.x + threadIdx.x] = x0 + x1 + x2 + x3 + x4 + x5 + x6 + x7;
}
// v_fma_f64 (the double-precision evaluation of the libm replicas)
__global__ void __launch_bounds__(256) fma_f64_chains(double* out, int iters, double a, double b) {
    double x0 = threadIdx.x, x1 = x0 + 1, x2 = x0 + 2, x3 = x0 + 3, x4 = x0 + 4, x5 = x0 + 5, x6 = x0 + 6, x7 = x0 + 7;
#define OP_F64(x) x = __builtin_fma(x, a, b)
    PROBE_BODY(OP_F64)
    out[blockIdx.x * blockDim.x + threadIdx.x] = x0 + x1 + x2 + x3 + x4 + x5 + x6 + x7;
}
// v_rcp_f32 (the transcendental unit: reciprocal and square-root estimates)
__global__ void __launch_bounds__(256) trans_f32_chains(float* out, int iters) {
    float x0 = threadIdx.x + 1, x1 = x0 + 1, x2 = x0 + 2, x3 = x0 + 3, x4 = x0 + 4, x5 = x0 + 5, x6 = x0 + 6, x7 = x0 + 7;
#define OP_RCP(x) x = __builtin_amdgcn_rcpf(x)
    PROBE_BODY(OP_RCP)
    out[blockIdx.x * blockDim.x + threadIdx.x] = x0 + x1 + x2 + x3 + x4 + x5 + x6 + x7;
}

__device__ __forceinline__ uint32_t mix(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
    return x;
}

__global__ void __launch_bounds__(256) spread_atomics(float* buf, uint32_t n, uint32_t iters) {
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    for (uint32_t k = 0; k < iters; k++) unsafeAtomicAdd(buf + mix(tid * 2654435761u + k) % n, 1.0f);
}

// the same hashed atomics with the non-temporal or system-scope cache-policy bits (no return value)
template <int POLICY>
__global__ void __launch_bounds__(256) spread_atomics_policy(float* buf, uint32_t n, uint32_t iters) {
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    for (uint32_t k = 0; k < iters; k++) {
        float* p = buf + mix(tid * 2654435761u + k) % n;
        const float one = 1.0f;
        if (POLICY == 1) asm volatile("global_atomic_add_f32 %0, %1, off nt" : : "v"(p), "v"(one) : "memory");
        else asm volatile("global_atomic_add_f32 %0, %1, off sc1" : : "v"(p), "v"(one) : "memory");
    }
}

// Times each issue-rate probe: wave64 instructions of its class per second (the counters of the
// same dispatches, SQ_INSTS_VALU and its per-class parts, check the count).
static int run_valu_probes() {
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    const int blocks = 16384, threads = 256, iters = 1000;
    void* out;
    CK(hipMalloc(&out, (size_t)blocks * threads * 8));
    const double winstr = (double)blocks * threads / 64 * iters * 128;
    struct Probe {
        const char* name;
        int per_op;   // wave instructions of the class per chain step
        void (*launch)(void*, int, int, int);
    };
    const Probe probes[] = {
        {"fma_f32", 1, [](void* o, int b, int t, int it) { hipLaunchKernelGGL(fma_chains, dim3(b), dim3(t), 0, 0, (float*)o, it, 0.999f, 0.001f); }},
        {"add_f32", 1, [](void* o, int b, int t, int it) { hipLaunchKernelGGL(add_f32_chains, dim3(b), dim3(t), 0, 0, (float*)o, it, 0.5f); }},
        {"int32_add_xor", 2, [](void* o, int b, int t, int it) { hipLaunchKernelGGL(int32_chains, dim3(b), dim3(t), 0, 0, (uint32_t*)o, it, 0x9E3779B9u, 0x85EBCA6Bu); }},
        {"cmp_cndmask", 2, [](void* o, int b, int t, int it) { hipLaunchKernelGGL(select_chains, dim3(b), dim3(t), 0, 0, (float*)o, it, 1e30f, 0.5f); }},
        {"mov_b32", 1, [](void* o, int b, int t, int it) { hipLaunchKernelGGL(mov_chains, dim3(b), dim3(t), 0, 0, (float*)o, it); }},
        {"fma_f64", 1, [](void* o, int b, int t, int it) { hipLaunchKernelGGL(fma_f64_chains, dim3(b), dim3(t), 0, 0, (double*)o, it, 0.999, 0.001); }},
        {"rcp_f32", 1, [](void* o, int b, int t, int it) { hipLaunchKernelGGL(trans_f32_chains, dim3(b), dim3(t), 0, 0, (float*)o, it); }},
    };
    printf("{\"valu_probes\": {");
    bool first = true;
    for (const Probe& p : probes) {
        float ms = 0.f;
        for (int rep = 0; rep < 2; rep++) {            // the second launch is timed
            CK(hipEventRecord(e0));
            p.launch(out, blocks, threads, iters);
            CK(hipGetLastError());
            CK(hipEventRecord(e1));
            CK(hipEventSynchronize(e1));
            CK(hipEventElapsedTime(&ms, e0, e1));
        }
        const double n = winstr * p.per_op;
        printf("%s\"%s\": {\"ms\": %.4f, \"wave_instr\": %.6e, \"wave_instr_per_s\": %.4e}", first ? "" : ", ", p.name,
               ms, n, n / (ms * 1e-3));
        first = false;
    }
    printf("}}\n");
    CK(hipFree(out));
    return 0;
}

int main(int argc, char** argv) {
    if (argc > 1 && std::string(argv[1]) == "--valu") return run_valu_probes();
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    auto time_ms = [&](auto launch) {
        CK(hipEventRecord(e0));
        launch();
        CK(hipGetLastError());
        CK(hipEventRecord(e1));
        CK(hipEventSynchronize(e1));
        float ms = 0.f;
        CK(hipEventElapsedTime(&ms, e0, e1));
        return (double)ms;
    };

    // VALU: 32768 workgroups x 256 lanes x 2000 x 128 FMAs
    const int fblocks = 32768, fthreads = 256, fiters = 2000;
    float* fout;
    CK(hipMalloc(&fout, (size_t)fblocks * fthreads * 4));
    auto fma_run = [&] { hipLaunchKernelGGL(fma_chains, dim3(fblocks), dim3(fthreads), 0, 0, fout, fiters, 0.999f, 0.001f); };
    time_ms(fma_run);
    const double fms = time_ms(fma_run);
    const double winstr = (double)fblocks * fthreads / 64 * fiters * 128;

    // atomics: 4096 workgroups x 256 lanes x 256 adds
    const uint32_t n = 10u << 20, ablocks = 4096, athreads = 256, aiters = 256;
    const uint32_t nmax = 256u << 20;     // the footprint sweep goes up to 1 GiB of floats
    float* buf;
    CK(hipMalloc(&buf, (size_t)nmax * 4));
    CK(hipMemset(buf, 0, (size_t)nmax * 4));
    auto spread = [&] { hipLaunchKernelGGL(spread_atomics, dim3(ablocks), dim3(athreads), 0, 0, buf, n, aiters); };
    auto hot = [&] { hipLaunchKernelGGL(spread_atomics, dim3(ablocks), dim3(athreads), 0, 0, buf, 256u, aiters); };
    time_ms(spread);
    const double sms = time_ms(spread);
    const double hms = time_ms(hot);
    const double natom = (double)ablocks * athreads * aiters;
    const uint32_t fps[] = {16u << 10, 256u << 10, 1u << 20, 4u << 20, 40u << 20, 100u << 20, 256u << 20};
    double fp_rate[sizeof fps / sizeof fps[0]];
    for (size_t i = 0; i < sizeof fps / sizeof fps[0]; i++) {
        auto run = [&] { hipLaunchKernelGGL(spread_atomics, dim3(ablocks), dim3(athreads), 0, 0, buf, fps[i], aiters); };
        time_ms(run);
        fp_rate[i] = natom / (time_ms(run) * 1e-3);
    }

    hipDeviceProp_t prop;
    CK(hipGetDeviceProperties(&prop, 0));
    printf("{\"device\": \"%s\", \"cus\": %d, \"valu_wave_instr_per_s\": %.4e, \"valu_ms\": %.3f, "
           "\"spread_atomics_per_s\": %.4e, \"hot_atomics_per_s\": %.4e}\n",
           prop.gcnArchName, prop.multiProcessorCount, winstr / (fms * 1e-3), fms, natom / (sms * 1e-3),
           natom / (hms * 1e-3));
    auto nt = [&] { hipLaunchKernelGGL(spread_atomics_policy<1>, dim3(ablocks), dim3(athreads), 0, 0, buf, n, aiters); };
    auto sc1 = [&] { hipLaunchKernelGGL(spread_atomics_policy<2>, dim3(ablocks), dim3(athreads), 0, 0, buf, n, aiters); };
    time_ms(nt);
    const double nt_rate = natom / (time_ms(nt) * 1e-3);
    time_ms(sc1);
    const double sc1_rate = natom / (time_ms(sc1) * 1e-3);
    printf("{\"spread_atomics_nt_per_s\": %.4e, \"spread_atomics_sc1_per_s\": %.4e}\n", nt_rate, sc1_rate);
    printf("{\"spread_atomics_per_s_by_footprint_mib\": {");
    for (size_t i = 0; i < sizeof fps / sizeof fps[0]; i++)
        printf("%s\"%g\": %.4e", i ? ", " : "", fps[i] * 4.0 / (1 << 20), fp_rate[i]);
    printf("}}\n");
    CK(hipFree(fout));
    CK(hipFree(buf));
    return 0;
}
