// State lifetime, preparation, transfers, errors and timing for libaqs_hip.so.
//
// The reference keeps the state in an af::array (QSimulator::statevector_,
// include/quantum.h:737); here it is one hipMalloc'd HBM buffer of 2^n
// interleaved complex values per aqs_state.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <cstring>
#include <mutex>

#include "internal.h"

using namespace aqs_hip;

namespace aqs_hip {

static thread_local std::string g_error;

void set_error(const std::string& msg) { g_error = msg; }
void clear_error() { g_error.clear(); }

static std::mutex g_init_mutex;
static std::atomic<int> g_device{-1};
static std::atomic<int> g_cus{0};
static std::atomic<hipStream_t> g_stream{nullptr};

static void init_device(int device) {
    std::lock_guard<std::mutex> lock(g_init_mutex);
    if (g_device == device && g_stream) return;
    int count = 0;
    AQS_HIP(hipGetDeviceCount(&count));
    if (count <= 0) fail(AQS_ERR_DEVICE, "no HIP device visible");
    if (device < 0 || device >= count)
        fail(AQS_ERR_OUT_OF_RANGE, "device index " + std::to_string(device) + " out of range");
    AQS_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    AQS_HIP(hipGetDeviceProperties(&prop, device));
    if (!g_stream || g_device != device) {
        hipStream_t st;
        AQS_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        g_stream = st;
    }
    g_cus = prop.multiProcessorCount;
    g_device = device;
}

hipStream_t default_stream() {
    if (!g_stream) init_device(0);
    return g_stream;
}

int current_device() {
    if (g_device < 0) init_device(0);
    return g_device;
}

int device_cu_count() {
    if (g_device < 0) init_device(0);
    return g_cus;
}

void ensure_staging(aqs_state s, size_t bytes) {
    if (s->staging_bytes >= bytes) return;
    if (s->staging) {
        AQS_HIP(hipStreamSynchronize(s->stream));
        AQS_HIP(hipFree(s->staging));
        s->staging = nullptr;
        s->staging_bytes = 0;
    }
    size_t sz = 4096;
    while (sz < bytes) sz *= 2;
    AQS_HIP(hipMalloc(&s->staging, sz));
    s->staging_bytes = sz;
}

void ensure_pinned(aqs_state s, size_t bytes) {
    if (s->pinned_bytes >= bytes) return;
    if (s->pinned) {
        AQS_HIP(hipStreamSynchronize(s->stream));
        AQS_HIP(hipHostFree(s->pinned));
        s->pinned = nullptr;
        s->pinned_bytes = 0;
    }
    size_t sz = 4096;
    while (sz < bytes) sz *= 2;
    AQS_HIP(hipHostMalloc(&s->pinned, sz, hipHostMallocDefault));
    s->pinned_bytes = sz;
}

// ---- profiling -------------------------------------------------------------

static thread_local std::vector<hipEvent_t> t_event_pool;

static hipEvent_t take_event() {
    if (!t_event_pool.empty()) {
        hipEvent_t e = t_event_pool.back();
        t_event_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    AQS_HIP(hipEventCreate(&e));
    return e;
}

void prof_begin(aqs_state s, int slot, hipEvent_t* ev_start) {
    (void)slot;
    if (!s->profiling) return;
    *ev_start = take_event();
    AQS_HIP(hipEventRecord(*ev_start, s->stream));
}

void prof_end(aqs_state s, int slot, hipEvent_t ev_start) {
    if (!s->profiling || !ev_start) return;
    hipEvent_t stop = take_event();
    AQS_HIP(hipEventRecord(stop, s->stream));
    s->prof_pending.push_back(ProfileRec{ev_start, stop, slot});
    if (s->prof_pending.size() >= 4096) prof_collect(s);
}

void prof_collect(aqs_state s) {
    if (s->prof_pending.empty()) return;
    AQS_HIP(hipStreamSynchronize(s->stream));
    for (auto& r : s->prof_pending) {
        float ms = 0.f;
        AQS_HIP(hipEventElapsedTime(&ms, r.start, r.stop));
        s->prof_ms[r.slot] += ms;
        s->prof_count[r.slot] += 1;
        t_event_pool.push_back(r.start);
        t_event_pool.push_back(r.stop);
    }
    s->prof_pending.clear();
}

}  // namespace aqs_hip

// ---- kernels -----------------------------------------------------------------

namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(uint64_t units, int per_thread = 1) {
    uint64_t threads = (units + per_thread - 1) / per_thread;
    uint64_t blocks = (threads + kBlock - 1) / kBlock;
    if (blocks > (1u << 20)) blocks = 1u << 20;
    if (blocks == 0) blocks = 1;
    return (unsigned)blocks;
}

template <typename T2>
__global__ void k_fill_zero(T2* __restrict__ a, uint64_t N) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    T2 z;
    z.x = 0;
    z.y = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += stride) a[i] = z;
}

template <typename T2>
__global__ void k_set_one(T2* __restrict__ a, uint64_t index) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        T2 v;
        v.x = 1;
        v.y = 0;
        a[index] = v;
    }
}

// Kronecker product of per-qubit 2-vectors (qubit 0 = most significant).
template <typename T2>
__global__ void k_product(T2* __restrict__ a, uint64_t N, int n, const double* __restrict__ qa) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += stride) {
        double re = 1.0, im = 0.0;
        for (int q = 0; q < n; ++q) {
            int b = (int)((i >> (n - 1 - q)) & 1ull);
            double xr = qa[4 * q + 2 * b], xi = qa[4 * q + 2 * b + 1];
            double nr = re * xr - im * xi;
            double ni = re * xi + im * xr;
            re = nr;
            im = ni;
        }
        T2 v;
        v.x = (decltype(v.x))re;
        v.y = (decltype(v.y))im;
        a[i] = v;
    }
}

__device__ inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__device__ inline double unit_uniform(uint64_t seed, uint64_t ctr) {
    uint64_t h = splitmix64(seed ^ splitmix64(ctr));
    return ((double)(h >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

// Box-Muller in FP64 on the counter pair (2i, 2i+1).
template <typename T2>
__global__ void k_gaussian(T2* __restrict__ a, uint64_t N, uint64_t seed) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += stride) {
        double u1 = unit_uniform(seed, 2 * i);
        double u2 = unit_uniform(seed, 2 * i + 1);
        double r = sqrt(-2.0 * log(u1));
        double ang = 6.283185307179586476925 * u2;
        T2 v;
        v.x = (decltype(v.x))(r * cos(ang));
        v.y = (decltype(v.y))(r * sin(ang));
        a[i] = v;
    }
}

// Plain streaming copy, 16 bytes per lane per access (also the calibration
// kernel for HBM traffic counters: it moves exactly 2 x bytes).  U float4s
// per lane, BS threads per block: one-wave blocks of one float4 per lane
// stream fastest (6.7 TB/s at 2 GiB against 5.8 TB/s for 256 x 4).
template <int U, int BS>
__global__ __launch_bounds__(BS) void k_copy16(float4* __restrict__ d, const float4* __restrict__ s, uint64_t n16) {
    const uint64_t base = (uint64_t)blockIdx.x * (BS * U) + threadIdx.x;
    float4 v[U];
#pragma unroll
    for (int j = 0; j < U; ++j) {
        const uint64_t i = base + (uint64_t)j * BS;
        if (i < n16) v[j] = ld_nt(&s[i]);
    }
#pragma unroll
    for (int j = 0; j < U; ++j) {
        const uint64_t i = base + (uint64_t)j * BS;
        if (i < n16) st_nt(v[j], &d[i]);
    }
}

template <typename TD, typename TS>
__global__ void k_convert(TD* __restrict__ d, const TS* __restrict__ s, uint64_t N) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += stride) {
        TS v = s[i];
        TD w;
        w.x = (decltype(w.x))v.x;
        w.y = (decltype(w.y))v.y;
        d[i] = w;
    }
}

void fill_zero(aqs_state s) {
    ProfScope ps(s, 0);
    if (s->precision == AQS_C64)
        hipLaunchKernelGGL(k_fill_zero<float2>, dim3(grid_for(s->N)), dim3(kBlock), 0, s->stream,
                           (float2*)s->amps, s->N);
    else
        hipLaunchKernelGGL(k_fill_zero<double2>, dim3(grid_for(s->N)), dim3(kBlock), 0, s->stream,
                           (double2*)s->amps, s->N);
    AQS_HIP(hipGetLastError());
}

aqs_state check_state(aqs_state s) {
    if (!s || !s->amps) fail(AQS_ERR_INVALID_ARGUMENT, "null state handle");
    return s;
}

}  // namespace

// ---- C ABI -------------------------------------------------------------------

namespace aqs_hip {
const char* last_error_cstr() { return g_error.c_str(); }
}  // namespace aqs_hip

extern "C" {

const char* aqs_last_error(void) { return aqs_hip::last_error_cstr(); }

aqs_status aqs_init(int device) {
    return guard([&] { aqs_hip::init_device(device); });
}

aqs_status aqs_device_info(char* name, size_t name_len, int* cu_count, size_t* hbm_bytes) {
    return guard([&] {
        hipDeviceProp_t prop;
        AQS_HIP(hipGetDeviceProperties(&prop, current_device()));
        if (name && name_len) {
            std::strncpy(name, prop.gcnArchName, name_len - 1);
            name[name_len - 1] = 0;
        }
        if (cu_count) *cu_count = prop.multiProcessorCount;
        if (hbm_bytes) *hbm_bytes = prop.totalGlobalMem;
    });
}

aqs_status aqs_state_create(int nqubits, int precision, aqs_state* out) {
    return guard([&] {
        if (!out) fail(AQS_ERR_INVALID_ARGUMENT, "null output handle");
        *out = nullptr;
        if (nqubits < 1) fail(AQS_ERR_INVALID_ARGUMENT, "Circuit must contain at least 1 qubit");
        if (nqubits > AQS_MAX_QUBITS)
            fail(AQS_ERR_INVALID_ARGUMENT, "Maximum qubit count supported is " + std::to_string(AQS_MAX_QUBITS));
        if (precision != AQS_C64 && precision != AQS_C128) fail(AQS_ERR_INVALID_ARGUMENT, "unknown precision");
        current_device();
        auto* s = new aqs_state_s();
        s->n = nqubits;
        s->precision = precision;
        s->N = 1ull << nqubits;
        s->stream = default_stream();
        hipError_t e = hipMalloc(&s->amps, s->N * s->amp_bytes());
        if (e != hipSuccess) {
            delete s;
            hip_check(e, "hipMalloc(state)");
        }
        // Scratch: reduction partials and the sampling block tables (three
        // 64-bit words per 1024 amplitudes), at least 64 KiB.
        size_t blocks = (size_t)((s->N + 1023) / 1024);
        s->scratch_bytes = std::max<size_t>(65536, (blocks + 1024) * 32);
        e = hipMalloc(&s->scratch, s->scratch_bytes);
        if (e != hipSuccess) {
            (void)hipFree(s->amps);
            delete s;
            hip_check(e, "hipMalloc(scratch)");
        }
        try {
            fill_zero(s);
            if (s->precision == AQS_C64)
                hipLaunchKernelGGL(k_set_one<float2>, dim3(1), dim3(1), 0, s->stream, (float2*)s->amps, 0ull);
            else
                hipLaunchKernelGGL(k_set_one<double2>, dim3(1), dim3(1), 0, s->stream, (double2*)s->amps, 0ull);
            AQS_HIP(hipGetLastError());
        } catch (...) {
            (void)hipFree(s->amps);
            (void)hipFree(s->scratch);
            delete s;
            throw;
        }
        *out = s;
    });
}

aqs_status aqs_state_destroy(aqs_state s) {
    return guard([&] {
        if (!s) return;
        (void)hipStreamSynchronize(s->stream);
        prof_collect(s);
        if (s->amps) (void)hipFree(s->amps);
        if (s->scratch) (void)hipFree(s->scratch);
        if (s->staging) (void)hipFree(s->staging);
        if (s->pinned) (void)hipHostFree(s->pinned);
        if (s->timer_start) (void)hipEventDestroy(s->timer_start);
        if (s->timer_stop) (void)hipEventDestroy(s->timer_stop);
        delete s;
    });
}

aqs_status aqs_state_copy(aqs_state dst, aqs_state src) {
    return guard([&] {
        check_state(dst);
        check_state(src);
        if (dst->n != src->n) fail(AQS_ERR_INVALID_ARGUMENT, "state copy between different qubit counts");
        if (src->stream != dst->stream) AQS_HIP(hipStreamSynchronize(src->stream));
        if (dst->precision == src->precision) {
            const uint64_t bytes = dst->N * dst->amp_bytes();
            if (bytes % 16 == 0) {
                const uint64_t n16 = bytes / 16;
                ProfScope ps(dst, 0);
                if (n16 <= (1ull << 31))
                    hipLaunchKernelGGL((k_copy16<1, 64>), dim3((unsigned)((n16 + 63) / 64)), dim3(64), 0,
                                       dst->stream, (float4*)dst->amps, (const float4*)src->amps, n16);
                else
                    hipLaunchKernelGGL((k_copy16<4, 256>), dim3((unsigned)((n16 + 1023) / 1024)), dim3(256), 0,
                                       dst->stream, (float4*)dst->amps, (const float4*)src->amps, n16);
            } else {
                AQS_HIP(hipMemcpyAsync(dst->amps, src->amps, bytes, hipMemcpyDeviceToDevice, dst->stream));
            }
        } else if (dst->precision == AQS_C64) {
            hipLaunchKernelGGL((k_convert<float2, double2>), dim3(grid_for(dst->N)), dim3(kBlock), 0, dst->stream,
                               (float2*)dst->amps, (const double2*)src->amps, dst->N);
        } else {
            hipLaunchKernelGGL((k_convert<double2, float2>), dim3(grid_for(dst->N)), dim3(kBlock), 0, dst->stream,
                               (double2*)dst->amps, (const float2*)src->amps, dst->N);
        }
        AQS_HIP(hipGetLastError());
    });
}

aqs_status aqs_state_qubits(aqs_state s, int* nqubits) {
    return guard([&] {
        check_state(s);
        if (nqubits) *nqubits = s->n;
    });
}

aqs_status aqs_state_precision(aqs_state s, int* precision) {
    return guard([&] {
        check_state(s);
        if (precision) *precision = s->precision;
    });
}

aqs_status aqs_state_device_ptr(aqs_state s, void** ptr) {
    return guard([&] {
        check_state(s);
        if (ptr) *ptr = s->amps;
    });
}

aqs_status aqs_state_set_stream(aqs_state s, void* hip_stream) {
    return guard([&] {
        check_state(s);
        AQS_HIP(hipStreamSynchronize(s->stream));
        prof_collect(s);
        s->stream = hip_stream ? (hipStream_t)hip_stream : default_stream();
    });
}

aqs_status aqs_state_sync(aqs_state s) {
    return guard([&] {
        check_state(s);
        AQS_HIP(hipStreamSynchronize(s->stream));
    });
}

aqs_status aqs_state_set_basis(aqs_state s, uint64_t index) {
    return guard([&] {
        check_state(s);
        if (index >= s->N) fail(AQS_ERR_OUT_OF_RANGE, "basis index out of range");
        fill_zero(s);
        if (s->precision == AQS_C64)
            hipLaunchKernelGGL(k_set_one<float2>, dim3(1), dim3(1), 0, s->stream, (float2*)s->amps, index);
        else
            hipLaunchKernelGGL(k_set_one<double2>, dim3(1), dim3(1), 0, s->stream, (double2*)s->amps, index);
        AQS_HIP(hipGetLastError());
    });
}

aqs_status aqs_state_product(aqs_state s, const double* qubit_amps) {
    return guard([&] {
        check_state(s);
        if (!qubit_amps) fail(AQS_ERR_INVALID_ARGUMENT, "null qubit amplitudes");
        size_t bytes = sizeof(double) * 4 * (size_t)s->n;
        ensure_staging(s, bytes);
        AQS_HIP(hipMemcpyAsync(s->staging, qubit_amps, bytes, hipMemcpyHostToDevice, s->stream));
        ProfScope ps(s, 0);
        if (s->precision == AQS_C64)
            hipLaunchKernelGGL(k_product<float2>, dim3(grid_for(s->N)), dim3(kBlock), 0, s->stream,
                               (float2*)s->amps, s->N, s->n, (const double*)s->staging);
        else
            hipLaunchKernelGGL(k_product<double2>, dim3(grid_for(s->N)), dim3(kBlock), 0, s->stream,
                               (double2*)s->amps, s->N, s->n, (const double*)s->staging);
        AQS_HIP(hipGetLastError());
        // The staging buffer is reused by the next call: finish the copy first.
        AQS_HIP(hipStreamSynchronize(s->stream));
    });
}

aqs_status aqs_state_gaussian(aqs_state s, uint64_t seed) {
    return guard([&] {
        check_state(s);
        {
            ProfScope ps(s, 0);
            if (s->precision == AQS_C64)
                hipLaunchKernelGGL(k_gaussian<float2>, dim3(grid_for(s->N)), dim3(kBlock), 0, s->stream,
                                   (float2*)s->amps, s->N, seed);
            else
                hipLaunchKernelGGL(k_gaussian<double2>, dim3(grid_for(s->N)), dim3(kBlock), 0, s->stream,
                                   (double2*)s->amps, s->N, seed);
            AQS_HIP(hipGetLastError());
        }
        double n2 = reduce_norm2(s);
        scale_state(s, 1.0 / std::sqrt(n2));
    });
}

aqs_status aqs_state_upload(aqs_state s, const void* host, int normalize) {
    return guard([&] {
        check_state(s);
        if (!host) fail(AQS_ERR_INVALID_ARGUMENT, "null host buffer");
        AQS_HIP(hipMemcpyAsync(s->amps, host, s->N * s->amp_bytes(), hipMemcpyHostToDevice, s->stream));
        AQS_HIP(hipStreamSynchronize(s->stream));
        if (normalize) {
            double n2 = reduce_norm2(s);
            if (!(n2 > 0.0)) fail(AQS_ERR_INVALID_ARGUMENT, "Cannot have a null statevector");
            scale_state(s, 1.0 / std::sqrt(n2));
        }
    });
}

aqs_status aqs_state_download(aqs_state s, void* host) {
    return guard([&] {
        check_state(s);
        if (!host) fail(AQS_ERR_INVALID_ARGUMENT, "null host buffer");
        AQS_HIP(hipMemcpyAsync(host, s->amps, s->N * s->amp_bytes(), hipMemcpyDeviceToHost, s->stream));
        AQS_HIP(hipStreamSynchronize(s->stream));
    });
}

aqs_status aqs_profile_enable(aqs_state s, int enable) {
    return guard([&] {
        check_state(s);
        prof_collect(s);
        s->profiling = enable != 0;
        for (int i = 0; i < kProfileSlots; ++i) {
            s->prof_ms[i] = 0.0;
            s->prof_count[i] = 0;
        }
    });
}

aqs_status aqs_profile_read(aqs_state s, double* ms_by_kind, uint64_t* launches_by_kind) {
    return guard([&] {
        check_state(s);
        prof_collect(s);
        for (int i = 0; i < kProfileSlots; ++i) {
            if (ms_by_kind) ms_by_kind[i] = s->prof_ms[i];
            if (launches_by_kind) launches_by_kind[i] = s->prof_count[i];
        }
    });
}

aqs_status aqs_timer_start(aqs_state s) {
    return guard([&] {
        check_state(s);
        if (!s->timer_start) AQS_HIP(hipEventCreate(&s->timer_start));
        if (!s->timer_stop) AQS_HIP(hipEventCreate(&s->timer_stop));
        AQS_HIP(hipEventRecord(s->timer_start, s->stream));
    });
}

aqs_status aqs_timer_stop(aqs_state s, double* ms) {
    return guard([&] {
        check_state(s);
        if (!s->timer_start) fail(AQS_ERR_RUNTIME, "timer not started");
        AQS_HIP(hipEventRecord(s->timer_stop, s->stream));
        AQS_HIP(hipEventSynchronize(s->timer_stop));
        float t = 0.f;
        AQS_HIP(hipEventElapsedTime(&t, s->timer_start, s->timer_stop));
        if (ms) *ms = t;
    });
}

}  // extern "C"
