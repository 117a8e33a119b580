// HBM bandwidth probe for the measurement side of the engine (bench.py):
// streaming kernels with a chosen number of read and written float arrays,
// 16 B per lane per access, with or without non-temporal hints.  They give
// the bandwidth a streaming kernel actually reaches on this device, for
// read-only, write-only, copy and the read/write mixes of the update kernels
// (a half step reads about 2.3 bytes for every byte it writes).
//
// C ABI (no counterpart in the reference; measurement only):
//   int ems_hbm_probe(int reads, int writes, int nontemporal,
//                     unsigned long long bytes_per_array, int reps,
//                     double* gbs, double* ms_per_launch)
// reads in [0, 8], writes in [0, 4], reads + writes >= 1.  Bytes counted per
// launch: (reads + writes) * bytes_per_array.  Returns 0, or a negative code
// with ems_hbm_probe_error() describing the failure.
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

constexpr unsigned kBlock = 256;
constexpr unsigned kUnroll = 4;      // 16-byte elements per lane

typedef float f4 __attribute__((ext_vector_type(4)));

// array pointers travel by value in the kernel arguments, so every access is
// a global (not flat) memory instruction
struct Arrays {
    const f4* src[8];
    f4* dst[4];
};

template <int R, int W, bool NT>
__global__ __launch_bounds__(kBlock) void k_stream(const Arrays a, unsigned long long n4, f4* __restrict__ sink)
{
    const unsigned long long base = (unsigned long long)blockIdx.x * kBlock * kUnroll + threadIdx.x;
    f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (unsigned u = 0; u < kUnroll; ++u) {
        const unsigned long long i = base + (unsigned long long)u * kBlock;
        if (i >= n4) break;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            f4 v;
            if (NT) v = __builtin_nontemporal_load(a.src[r] + i);
            else v = a.src[r][i];
            acc += v;
        }
        if (R == 0) acc = f4{(float)i, 1.f, 2.f, 3.f};
#pragma unroll
        for (int w = 0; w < W; ++w) {
            if (NT) __builtin_nontemporal_store(acc, a.dst[w] + i);
            else a.dst[w][i] = acc;
        }
    }
    // keeps the loads of a read-only launch alive; the inputs are zero
    if (W == 0 && acc.x == 1.0e30f) sink[threadIdx.x] = acc;
}

template <int R, int W>
hipError_t launch(bool nt, unsigned blocks, const Arrays& a, unsigned long long n4, f4* sink)
{
    if (nt) k_stream<R, W, true><<<blocks, kBlock>>>(a, n4, sink);
    else k_stream<R, W, false><<<blocks, kBlock>>>(a, n4, sink);
    return hipGetLastError();
}

using LaunchFn = hipError_t (*)(bool, unsigned, const Arrays&, unsigned long long, f4*);

template <int R>
LaunchFn pick_w(int w)
{
    switch (w) {
    case 0: return launch<R, 0>;
    case 1: return launch<R, 1>;
    case 2: return launch<R, 2>;
    case 3: return launch<R, 3>;
    case 4: return launch<R, 4>;
    }
    return nullptr;
}

LaunchFn pick(int r, int w)
{
    switch (r) {
    case 0: return pick_w<0>(w);
    case 1: return pick_w<1>(w);
    case 2: return pick_w<2>(w);
    case 3: return pick_w<3>(w);
    case 4: return pick_w<4>(w);
    case 5: return pick_w<5>(w);
    case 6: return pick_w<6>(w);
    case 7: return pick_w<7>(w);
    case 8: return pick_w<8>(w);
    }
    return nullptr;
}

int fail(const std::string& m)
{
    g_err = m;
    return -1;
}

}  // namespace

extern "C" {

const char* ems_hbm_probe_error(void) { return g_err.c_str(); }

int ems_hbm_probe(int reads, int writes, int nontemporal, unsigned long long bytes_per_array, int reps, double* gbs,
                  double* ms_per_launch)
{
    if (reads < 0 || reads > 8 || writes < 0 || writes > 4 || reads + writes < 1 || reps < 1 || !gbs ||
        !ms_per_launch)
        return fail("ems_hbm_probe: reads in [0,8], writes in [0,4], at least one array, reps >= 1");
    const unsigned long long n4 = bytes_per_array / 16;
    if (n4 == 0) return fail("ems_hbm_probe: bytes_per_array below 16");
    const unsigned long long per_block = (unsigned long long)kBlock * kUnroll;
    const unsigned long long nb = (n4 + per_block - 1) / per_block;
    if (nb > 0x7fffffffull) return fail("ems_hbm_probe: array too large");
    LaunchFn fn = pick(reads, writes);
    std::vector<f4*> bufs(reads + writes, nullptr);
    f4* sink = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    auto chk = [&](hipError_t e, const char* what) {
        if (e != hipSuccess && rc == 0) rc = fail(std::string(what) + ": " + hipGetErrorString(e));
        return e == hipSuccess;
    };
    do {
        bool ok = true;
        for (auto& b : bufs) {
            ok = chk(hipMalloc(&b, n4 * 16), "hipMalloc") && chk(hipMemset(b, 0, n4 * 16), "hipMemset");
            if (!ok) break;
        }
        if (!ok) break;
        if (!chk(hipMalloc(&sink, sizeof(f4) * kBlock), "hipMalloc")) break;
        if (!chk(hipEventCreate(&e0), "hipEventCreate") || !chk(hipEventCreate(&e1), "hipEventCreate")) break;
        Arrays a{};
        for (int r = 0; r < reads; ++r) a.src[r] = bufs[r];
        for (int w = 0; w < writes; ++w) a.dst[w] = bufs[reads + w];
        // two warm-up launches
        for (int i = 0; i < 2 && ok; ++i) ok = chk(fn(nontemporal != 0, (unsigned)nb, a, n4, sink), "launch");
        if (!ok || !chk(hipDeviceSynchronize(), "hipDeviceSynchronize")) break;
        chk(hipEventRecord(e0, 0), "hipEventRecord");
        for (int i = 0; i < reps && ok; ++i) ok = chk(fn(nontemporal != 0, (unsigned)nb, a, n4, sink), "launch");
        chk(hipEventRecord(e1, 0), "hipEventRecord");
        if (!chk(hipEventSynchronize(e1), "hipEventSynchronize") || rc) break;
        float ms = 0.f;
        if (!chk(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime")) break;
        *ms_per_launch = ms / reps;
        *gbs = (double)(reads + writes) * (double)(n4 * 16) / (*ms_per_launch * 1e-3) / 1e9;
    } while (false);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (sink) (void)hipFree(sink);
    for (auto b : bufs)
        if (b) (void)hipFree(b);
    return rc;
}

}  // extern "C"
