// Internal declarations shared by the libaqs_hip.so translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "aqs_hip.h"

namespace aqs_hip {

// Thread-local error message behind aqs_last_error().
void set_error(const std::string& msg);
void clear_error();

struct Status {
    aqs_status code;
};

// Throwable carrier used inside the library; converted to aqs_status at the
// C boundary by guard().
struct Error {
    aqs_status code;
    std::string msg;
};

[[noreturn]] inline void fail(aqs_status code, const std::string& msg) { throw Error{code, msg}; }

inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        if (e == hipErrorOutOfMemory) fail(AQS_ERR_NO_MEMORY, std::string(what) + ": out of device memory");
        fail(AQS_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
    }
}

#define AQS_HIP(call) ::aqs_hip::hip_check((call), #call)

template <typename F>
aqs_status guard(F&& f) {
    try {
        clear_error();
        f();
        return AQS_OK;
    } catch (const Error& e) {
        set_error(e.msg);
        return e.code;
    } catch (const std::bad_alloc&) {
        set_error("host allocation failed");
        return AQS_ERR_NO_MEMORY;
    } catch (const std::exception& e) {
        set_error(e.what());
        return AQS_ERR_RUNTIME;
    }
}

// Profile slots: index = aqs_op_kind (1..6) of a plain pass, 7 = MATK on
// MFMA, 8 = diagonal table, 9 = bit permutation, 10 = tile pass, 0 =
// everything else (state preparation, measurement).
constexpr int kProfileSlots = AQS_PROFILE_SLOTS;
constexpr int kSlotMfma = 7;
constexpr int kSlotDiagTable = 8;
constexpr int kSlotPermute = 9;
constexpr int kSlotTile = 10;
constexpr int kSlotSmall = 11;

struct ProfileRec {
    hipEvent_t start, stop;
    int slot;
};

}  // namespace aqs_hip

struct aqs_state_s {
    int n = 0;                 // qubits
    int precision = AQS_C64;   // aqs_precision
    uint64_t N = 0;            // 2^n amplitudes
    void* amps = nullptr;      // device: N x (float2 | double2)
    hipStream_t stream = nullptr;
    bool own_stream = false;

    // Device scratch for reductions and sampling (allocated with the state).
    void* scratch = nullptr;
    size_t scratch_bytes = 0;
    // Device staging for small uploads (gate matrices, uniforms).
    void* staging = nullptr;
    size_t staging_bytes = 0;
    // Pinned host buffer for small results.
    void* pinned = nullptr;
    size_t pinned_bytes = 0;

    // Profiling.
    bool profiling = false;
    std::vector<aqs_hip::ProfileRec> prof_pending;
    double prof_ms[aqs_hip::kProfileSlots] = {};
    uint64_t prof_count[aqs_hip::kProfileSlots] = {};
    hipEvent_t timer_start = nullptr, timer_stop = nullptr;

    size_t amp_bytes() const { return precision == AQS_C64 ? 8 : 16; }
};

namespace aqs_hip {

// Library stream for states without an external stream.
hipStream_t default_stream();
int current_device();
// Compute units of the current device (read once at initialisation).
int device_cu_count();

// Make sure the state's staging buffer holds at least `bytes`.
void ensure_staging(aqs_state s, size_t bytes);
void ensure_pinned(aqs_state s, size_t bytes);

// Profiling brackets around one kernel launch.
void prof_begin(aqs_state s, int slot, hipEvent_t* ev_start);
void prof_end(aqs_state s, int slot, hipEvent_t ev_start);
void prof_collect(aqs_state s);

struct ProfScope {
    aqs_state s;
    int slot;
    hipEvent_t ev = nullptr;
    ProfScope(aqs_state st, int sl) : s(st), slot(sl) { prof_begin(s, slot, &ev); }
    ~ProfScope() { prof_end(s, slot, ev); }
};

// Gate dispatch (gates.hip / mfma.hip).
void apply_op(aqs_state s, const aqs_op& op);
// A diagonal table pass (anchor, chunks, required bits as amplitude bit
// positions; dev_tab = its float2 table on the device) to apply in the
// epilogue of an MFMA block.
struct DiagEpilogueDesc {
    std::vector<int> required, anchor;
    std::vector<std::vector<int>> chunks;
    const void* dev_tab = nullptr;
};

// k-qubit dense block on MFMA (k = 4 or 5, complex64); returns false when
// the op shape is not supported by the MFMA path, or when `epi` is given
// and its anchor bits do not lie in the block's tile.  `epi` is applied
// after the block, or before it when `before` is set.  `beta` (a bit
// involution over amplitude bit positions) is applied after the block and
// any table before it, when it maps the block's tile bits onto themselves
// and no table follows; otherwise the call returns false.
bool apply_matk_mfma(aqs_state s, const aqs_op& op, const float* dev_apack, const DiagEpilogueDesc* epi = nullptr,
                     bool before = false, const std::vector<int>* beta = nullptr);
// MATK with the matrix already on the device (FP32 for complex64, FP64 for
// complex128, plain row-major).
void apply_matk_device(aqs_state s, const aqs_op& op, const void* dev_u);
// Pack an FP64 complex matrix (2^k x 2^k) into the MFMA operand layout.
std::vector<float> pack_mfma_operand(int k, const double* u);

// Diagonal table pass (diag.hip): bit positions in amplitude-index space;
// dev_tab = [2^|anchor|][chunks][32] complex in the state's precision.
void apply_diag_table(aqs_state s, const std::vector<int>& required, const std::vector<int>& anchor,
                      const std::vector<std::vector<int>>& chunks, const void* dev_tab);

// In-place index-bit permutation (perm.hip); beta: involution of bit positions.
void apply_bit_involution(aqs_state s, const std::vector<int>& beta);

// Validation shared by apply and fuse.
void validate_op(const aqs_op& op, int nqubits);

// Reductions (measure.hip).
double reduce_norm2(aqs_state s);
void scale_state(aqs_state s, double factor);

// Device bytes that queue_sample_norm needs as scratch for a complex64 state
// of 2^n amplitudes (sampling tables and reduction partials).
size_t sample_norm_scratch_bytes(int n);
// Queues on `stream`, without waiting: `shots` indices sampled from the
// complex64 state `amps` of 2^n amplitudes (uniforms du and indices dout on
// the device; the exact fixed-point CDF of aqs_sample) and its FP64 norm^2
// into *dnorm (device).
void queue_sample_norm(const void* amps, int n, void* scratch, const double* du, uint32_t shots, uint64_t* dout,
                       double* dnorm, hipStream_t stream);

// Streaming (non-temporal) 8- and 16-byte accesses: a gate pass touches each
// amplitude once, so the lines are not kept in the caches.
typedef float aqs_f2v __attribute__((ext_vector_type(2)));
typedef float aqs_f4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float4 ld_nt(const float4* p) {
    aqs_f4v v = __builtin_nontemporal_load((const aqs_f4v*)p);
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ float2 ld_nt(const float2* p) {
    aqs_f2v v = __builtin_nontemporal_load((const aqs_f2v*)p);
    return make_float2(v.x, v.y);
}
__device__ __forceinline__ void st_nt(float4 v, float4* p) {
    aqs_f4v w = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(w, (aqs_f4v*)p);
}
__device__ __forceinline__ void st_nt(float2 v, float2* p) {
    aqs_f2v w = {v.x, v.y};
    __builtin_nontemporal_store(w, (aqs_f2v*)p);
}

inline uint64_t qbit(int n, int q) { return 1ull << (uint64_t)(n - 1 - q); }
inline int qpos(int n, int q) { return n - 1 - q; }

}  // namespace aqs_hip
