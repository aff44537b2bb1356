// pw_hbm_probe -- measured HBM bandwidth of plain streaming kernels on this GPU.
//
// The step kernels are HBM-bound, so the rate that a kernel of the simplest
// possible shape reaches is the practical roof they are measured against, next
// to the datasheet figure.  Each pattern below streams separate 256 MiB arrays
// with 16-byte loads and stores in a grid-stride loop:
//   copy         1 read : 1 write
//   stream 2:1   2 reads : 1 write
//   stream 6:1   6 reads : 1 write
//   stream 12:1  12 reads : 1 write  (the step kernels' read/write mix and
//                                     number of concurrent streams)
//   read 12      12 reads, one value per thread written at the end
//   step-like    the velocity kernel's mix with 4-byte accesses: 17 arrays read
//                (13 wavefields + 4 coefficients), 6 of them rewritten in place
//   step-like, interleaved  the same accesses with the 17 arrays stored as one:
//                each 64-value row segment of array k at (row * 17 + k) * 64
//   step-like, out-of-place  17 arrays read, the 6 results written to 6 other
//                arrays instead of in place
//   stress-like  the stress kernel's FP32 mix: 18 arrays read (6 velocities,
//                7 stresses, 5 coefficients), the 7 stresses rewritten in place
//   comp16-like  the velocity kernel's compensated-FP16 mix: 13 fields of 2-byte
//                values and 1-byte residuals (6 rewritten in place) and 15
//                2-byte coefficient arrays, 1 or 2 cells per thread
// Rates count the bytes each pattern must move (every array once), timed
// with HIP events over repeated launches, best over a few grid sizes.
//
// Output: one line per pattern, "pattern blocks GB/s", then a JSON summary.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>

#define CHECK(x)                                                                                  \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess) {                                                                   \
            std::fprintf(stderr, "HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
            std::exit(1);                                                                         \
        }                                                                                         \
    } while (0)

namespace {

constexpr int kMaxIn = 12;
constexpr int kThreads = 256;
struct Inputs {
    const float4* a[kMaxIn];
};

__global__ __launch_bounds__(kThreads) void k_fill(float4* p, size_t n4, float v)
{
    for (size_t i = blockIdx.x * (size_t)kThreads + threadIdx.x; i < n4; i += (size_t)gridDim.x * kThreads)
        p[i] = make_float4(v, v, v, v);
}

// R reads : 1 write (R = 1 is a copy)
template <int R>
__global__ __launch_bounds__(kThreads) void k_stream(Inputs in, float4* __restrict__ out, size_t n4)
{
    for (size_t i = blockIdx.x * (size_t)kThreads + threadIdx.x; i < n4; i += (size_t)gridDim.x * kThreads) {
        float4 s = in.a[0][i];
#pragma unroll
        for (int r = 1; r < R; ++r) {
            const float4 v = in.a[r][i];
            s.x += v.x; s.y += v.y; s.z += v.z; s.w += v.w;
        }
        out[i] = s;
    }
}

// R reads, one float per thread written (sums keep the loads live)
template <int R>
__global__ __launch_bounds__(kThreads) void k_read(Inputs in, float* __restrict__ out, size_t n4)
{
    float acc = 0.0f;
    for (size_t i = blockIdx.x * (size_t)kThreads + threadIdx.x; i < n4; i += (size_t)gridDim.x * kThreads) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const float4 v = in.a[r][i];
            acc += v.x + v.y + v.z + v.w;
        }
    }
    out[blockIdx.x * (size_t)kThreads + threadIdx.x] = acc;
}

// 17 reads of separate arrays, the first 6 rewritten in place, one float per thread
struct StepLike {
    float* rw[6];
    const float* ro[11];
};
__global__ __launch_bounds__(kThreads) void k_step_like(StepLike a, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)kThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kThreads) {
        float acc = 0.0f;
#pragma unroll
        for (int k = 0; k < 11; ++k) acc += a.ro[k][i];
#pragma unroll
        for (int k = 0; k < 6; ++k) a.rw[k][i] = a.rw[k][i] * 0.5f + acc * 1e-3f;
    }
}

// 17 arrays read, 6 results written to separate arrays
struct StepLikeOut {
    const float* ro[17];
    float* wo[6];
};
__global__ __launch_bounds__(kThreads) void k_step_like_out(StepLikeOut a, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)kThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kThreads) {
        float acc = 0.0f;
#pragma unroll
        for (int k = 6; k < 17; ++k) acc += a.ro[k][i];
#pragma unroll
        for (int k = 0; k < 6; ++k) a.wo[k][i] = a.ro[k][i] * 0.5f + acc * 1e-3f;
    }
}

// 18 reads of separate arrays, the first 7 rewritten in place
struct StressLike {
    float* rw[7];
    const float* ro[11];
};
__global__ __launch_bounds__(kThreads) void k_stress_like(StressLike a, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)kThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kThreads) {
        float acc = 0.0f;
#pragma unroll
        for (int k = 0; k < 11; ++k) acc += a.ro[k][i];
#pragma unroll
        for (int k = 0; k < 7; ++k) a.rw[k][i] = a.rw[k][i] * 0.5f + acc * 1e-3f;
    }
}

// compensated-FP16 mix; N consecutive cells per thread (one 2N- / N-byte access)
template <class T, int N> struct alignas(sizeof(T) * N) Vec { T v[N]; };
struct Comp16Like {
    uint16_t* hi[13];
    int8_t* lo[13];
    const uint16_t* co[15];
};
template <int N>
__global__ __launch_bounds__(kThreads) void k_comp16_like(Comp16Like a, size_t n)
{
    using H = Vec<uint16_t, N>;
    using L = Vec<int8_t, N>;
    for (size_t c = blockIdx.x * (size_t)kThreads + threadIdx.x; c * N < n; c += (size_t)gridDim.x * kThreads) {
        const size_t i = c * N;
        unsigned acc[N] = {};
#pragma unroll
        for (int k = 6; k < 13; ++k) {
            const H h = *(const H*)(a.hi[k] + i);
            const L l = *(const L*)(a.lo[k] + i);
#pragma unroll
            for (int j = 0; j < N; ++j) acc[j] += h.v[j] + (unsigned)l.v[j];
        }
#pragma unroll
        for (int k = 0; k < 15; ++k) {
            const H h = *(const H*)(a.co[k] + i);
#pragma unroll
            for (int j = 0; j < N; ++j) acc[j] += h.v[j];
        }
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            H h = *(const H*)(a.hi[k] + i);
            L l = *(const L*)(a.lo[k] + i);
#pragma unroll
            for (int j = 0; j < N; ++j) {
                h.v[j] = (uint16_t)(h.v[j] + acc[j]);
                l.v[j] = (int8_t)(l.v[j] + (int)acc[j]);
            }
            *(H*)(a.hi[k] + i) = h;
            *(L*)(a.lo[k] + i) = l;
        }
    }
}

// the same 17 arrays interleaved by 64-value row segments in one buffer
__global__ __launch_bounds__(kThreads) void k_step_like_aosoa(float* __restrict__ buf, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)kThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kThreads) {
        const size_t base = (i / 64) * 17 * 64 + (i % 64);
        float acc = 0.0f;
#pragma unroll
        for (int k = 6; k < 17; ++k) acc += buf[base + (size_t)k * 64];
#pragma unroll
        for (int k = 0; k < 6; ++k) buf[base + (size_t)k * 64] = buf[base + (size_t)k * 64] * 0.5f + acc * 1e-3f;
    }
}

template <class F>
double best_rate(F launch, double bytes, int iters, const int* grids, int ngrids, int* best_grid)
{
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    double best = 0.0;
    for (int g = 0; g < ngrids; ++g) {
        launch(grids[g]);  // warm-up
        CHECK(hipGetLastError());
        CHECK(hipEventRecord(e0));
        for (int it = 0; it < iters; ++it) launch(grids[g]);
        CHECK(hipEventRecord(e1));
        CHECK(hipEventSynchronize(e1));
        float ms = 0.0f;
        CHECK(hipEventElapsedTime(&ms, e0, e1));
        const double rate = bytes * iters / (ms * 1e-3) / 1e9;
        if (rate > best) { best = rate; *best_grid = grids[g]; }
    }
    CHECK(hipEventDestroy(e0));
    CHECK(hipEventDestroy(e1));
    return best;
}

}  // namespace

int main(int argc, char** argv)
{
    const int iters = argc > 1 ? std::atoi(argv[1]) : 20;
    const size_t n = (size_t)1 << 26;  // floats per array: 256 MiB
    const size_t n4 = n / 4;
    const double arr = (double)n * sizeof(float);
    const int grids[4] = {1024, 2048, 4096, 16384};
    const int ngrids = 4;

    float4* buf[kMaxIn + 1];
    for (int r = 0; r <= kMaxIn; ++r) {
        CHECK(hipMalloc(&buf[r], n * sizeof(float)));
        k_fill<<<4096, kThreads>>>(buf[r], n4, 1.0f + r);
    }
    float* red = nullptr;
    CHECK(hipMalloc(&red, (size_t)grids[ngrids - 1] * kThreads * sizeof(float)));
    CHECK(hipDeviceSynchronize());
    Inputs in;
    for (int r = 0; r < kMaxIn; ++r) in.a[r] = buf[r];
    float4* out = buf[kMaxIn];

    struct Row { const char* name; double rate; int grid; } rows[11];
    int g = 0;
    rows[0] = {"copy", best_rate([&](int b) { k_stream<1><<<b, kThreads>>>(in, out, n4); }, 2 * arr, iters, grids, ngrids, &g), g};
    rows[1] = {"stream_2r1w", best_rate([&](int b) { k_stream<2><<<b, kThreads>>>(in, out, n4); }, 3 * arr, iters, grids, ngrids, &g), g};
    rows[2] = {"stream_6r1w", best_rate([&](int b) { k_stream<6><<<b, kThreads>>>(in, out, n4); }, 7 * arr, iters, grids, ngrids, &g), g};
    rows[3] = {"stream_12r1w", best_rate([&](int b) { k_stream<12><<<b, kThreads>>>(in, out, n4); }, 13 * arr, iters, grids, ngrids, &g), g};
    rows[4] = {"read_12", best_rate([&](int b) { k_read<12><<<b, kThreads>>>(in, red, n4); }, 12 * arr, iters, grids, ngrids, &g), g};
    // the step-like mix over the first 12 buffers plus 5 more arrays of the same size
    float* extra[5];
    for (int r = 0; r < 5; ++r) {
        CHECK(hipMalloc(&extra[r], n * sizeof(float)));
        k_fill<<<4096, kThreads>>>((float4*)extra[r], n4, 1e-3f);
    }
    StepLike sl;
    for (int r = 0; r < 6; ++r) sl.rw[r] = (float*)buf[r];
    for (int r = 0; r < 6; ++r) sl.ro[r] = (const float*)buf[6 + r];
    for (int r = 0; r < 5; ++r) sl.ro[6 + r] = extra[r];
    rows[5] = {"step_like", best_rate([&](int b) { k_step_like<<<b, kThreads>>>(sl, n); }, 23 * arr, iters, grids, ngrids, &g), g};
    // out of place: the 17 inputs as above, the 6 outputs in buffers 12 and the 5 extras
    StepLikeOut so;
    for (int r = 0; r < 12; ++r) so.ro[r] = (const float*)buf[r];
    for (int r = 0; r < 5; ++r) so.ro[12 + r] = extra[r];
    float* outs[6];
    for (int r = 0; r < 6; ++r) CHECK(hipMalloc(&outs[r], n * sizeof(float)));
    for (int r = 0; r < 6; ++r) so.wo[r] = outs[r];
    rows[7] = {"step_like_out_of_place",
               best_rate([&](int b) { k_step_like_out<<<b, kThreads>>>(so, n); }, 23 * arr, iters, grids, ngrids, &g), g};
    for (int r = 0; r < 6; ++r) CHECK(hipFree(outs[r]));
    // the stress kernel's mix: buffers 0-6 rewritten, 7-11 and the 5 extras and one more read
    {
        float* more = nullptr;
        CHECK(hipMalloc(&more, n * sizeof(float)));
        k_fill<<<4096, kThreads>>>((float4*)more, n4, 1e-3f);
        StressLike st;
        for (int r = 0; r < 7; ++r) st.rw[r] = (float*)buf[r];
        for (int r = 0; r < 5; ++r) st.ro[r] = (const float*)buf[7 + r];
        for (int r = 0; r < 5; ++r) st.ro[5 + r] = extra[r];
        st.ro[10] = more;
        rows[10] = {"stress_like", best_rate([&](int b) { k_stress_like<<<b, kThreads>>>(st, n); }, 25 * arr, iters, grids, ngrids, &g), g};
        CHECK(hipFree(more));
    }

    for (int r = 0; r < 5; ++r) CHECK(hipFree(extra[r]));
    // the compensated-FP16 mix: 13 x 3 B + 15 x 2 B read, 6 x 3 B written per cell
    {
        Comp16Like c16;
        void* blocks[13 + 13 + 15];
        int nb = 0;
        for (int k = 0; k < 13; ++k) {
            CHECK(hipMalloc(&blocks[nb], n * 2)); c16.hi[k] = (uint16_t*)blocks[nb++];
            CHECK(hipMalloc(&blocks[nb], n)); c16.lo[k] = (int8_t*)blocks[nb++];
        }
        for (int k = 0; k < 15; ++k) { CHECK(hipMalloc(&blocks[nb], n * 2)); c16.co[k] = (const uint16_t*)blocks[nb++]; }
        for (int k = 0; k < nb; ++k) CHECK(hipMemset(blocks[k], 0, k < 26 && (k % 2) ? n : n * 2));
        CHECK(hipDeviceSynchronize());
        const double bytes = (13.0 * 3 + 15.0 * 2 + 6.0 * 3) * (double)n;
        rows[8] = {"comp16_like", best_rate([&](int b) { k_comp16_like<1><<<b, kThreads>>>(c16, n); }, bytes, iters, grids, ngrids, &g), g};
        rows[9] = {"comp16_like_2cells", best_rate([&](int b) { k_comp16_like<2><<<b, kThreads>>>(c16, n); }, bytes, iters, grids, ngrids, &g), g};
        for (int k = 0; k < nb; ++k) CHECK(hipFree(blocks[k]));
    }
    // the interleaved layout of the same 17 arrays (n is a multiple of 64)
    float* inter = nullptr;
    CHECK(hipMalloc(&inter, 17 * n * sizeof(float)));
    k_fill<<<4096, kThreads>>>((float4*)inter, 17 * n4, 1e-3f);
    rows[6] = {"step_like_interleaved",
               best_rate([&](int b) { k_step_like_aosoa<<<b, kThreads>>>(inter, n); }, 23 * arr, iters, grids, ngrids, &g), g};
    CHECK(hipFree(inter));
    // the 2:1 stream of arrays filled with 1 and 2 must give 3
    for (int r = 0; r < 3; ++r) k_fill<<<4096, kThreads>>>(buf[r], n4, 1.0f + r);
    k_stream<2><<<1024, kThreads>>>(in, out, n4);
    float4 h;
    CHECK(hipMemcpy(&h, out + n4 - 1, sizeof(h), hipMemcpyDeviceToHost));
    if (h.x != 3.0f || h.w != 3.0f) {
        std::fprintf(stderr, "probe check failed: %g\n", h.x);
        return 1;
    }
    for (const Row& r : rows) std::printf("%-14s %6d blocks %8.1f GB/s\n", r.name, r.grid, r.rate);
    std::printf("{");
    for (int i = 0; i < 11; ++i) std::printf("%s\"%s\": %.1f", i ? ", " : "", rows[i].name, rows[i].rate);
    std::printf(", \"unit\": \"GB/s\", \"array_bytes\": %.0f, \"iters\": %d}\n", arr, iters);
    for (int r = 0; r <= kMaxIn; ++r) CHECK(hipFree(buf[r]));
    CHECK(hipFree(red));
    return 0;
}
