// Bulk atomic operations on device values (src/stdgpu/atomic.cuh: atomic<T> and atomic_ref<T>).
//
// The reference applies one atomic operation per device thread (its tests run
// for_each_index(N, f) with f calling value.fetch_add(sequence[i]), tests/stdgpu/atomic.inc:766-883).
// The bulk form takes n operations at once: operation i applies operands[i] to element targets[i]
// (element 0 for every i when targets is NULL) and returns the value it replaced, as the device call
// returns it.  Each operation is atomic and the results are those of some order of the n operations,
// which is what n concurrent device calls give.
//
// When the operation composes (add, sub, min, max, and, or, xor, exchange, store on the integer
// types) and every live operation of a workgroup (1024 threads) names one element, the workgroup
// applies them all with one hardware atomic: a lane-order scan in each wavefront gives its combined
// operand and each lane's prefix, one thread combines the wavefronts' operands in wavefront order,
// applies the combination and hands each wavefront the value its first operation sees, and thread j's
// old value is that value combined with the operands of the live threads before j.  That is the result
// of applying the workgroup's operations one after another in thread order, so it is a valid order of
// the n operations.  A wavefront whose live operations name one element while the workgroup's do not
// does the same with one atomic of its own; the others apply one atomic per operation.  A counter that
// every thread of a launch hits then takes n / 1024 atomics instead of n.
#include <hip/hip_runtime.h>

#include <limits>
#include <new>
#include <type_traits>

#include <stdgpu_amd/atomic_device.hpp>
#include <stdgpu_amd/htab.h>

#include "wave.hpp"

using namespace htab_dev;

namespace
{

template <typename T, int OP>
constexpr bool kAggregates = !std::is_same_v<T, float> &&
                             (OP == HTAB_ATOMIC_FETCH_ADD || OP == HTAB_ATOMIC_FETCH_SUB ||
                              OP == HTAB_ATOMIC_FETCH_MIN || OP == HTAB_ATOMIC_FETCH_MAX ||
                              OP == HTAB_ATOMIC_FETCH_AND || OP == HTAB_ATOMIC_FETCH_OR ||
                              OP == HTAB_ATOMIC_FETCH_XOR || OP == HTAB_ATOMIC_EXCHANGE || OP == HTAB_ATOMIC_STORE);

// the operand that leaves a value unchanged, for the scan over lanes without an operation
template <typename T, int OP>
__device__ inline T
identity()
{
    if constexpr (OP == HTAB_ATOMIC_FETCH_AND) return static_cast<T>(~std::make_unsigned_t<T>(0));
    else if constexpr (OP == HTAB_ATOMIC_FETCH_MIN) return std::numeric_limits<T>::max();
    else if constexpr (OP == HTAB_ATOMIC_FETCH_MAX) return std::numeric_limits<T>::min();
    else return T(0);
}

// a then b, combined into one operand (sub combines its operands by addition)
template <typename T, int OP>
__device__ inline T
combine(T a, T b)
{
    using U = std::make_unsigned_t<T>;
    if constexpr (OP == HTAB_ATOMIC_FETCH_ADD || OP == HTAB_ATOMIC_FETCH_SUB)
        return static_cast<T>(static_cast<U>(a) + static_cast<U>(b));
    else if constexpr (OP == HTAB_ATOMIC_FETCH_MIN) return b < a ? b : a;
    else if constexpr (OP == HTAB_ATOMIC_FETCH_MAX) return a < b ? b : a;
    else if constexpr (OP == HTAB_ATOMIC_FETCH_AND) return a & b;
    else if constexpr (OP == HTAB_ATOMIC_FETCH_OR) return a | b;
    else return a ^ b;
}

// one operation on *p by one lane; ok = false for a compare-exchange that did not exchange
template <typename T, int OP>
__device__ inline T
apply_one(T* p, T arg, T expected, bool& ok)
{
    constexpr mem_order r = mem_order::relaxed;
    if constexpr (OP == HTAB_ATOMIC_LOAD) return atomic_load_op(p, r);
    else if constexpr (OP == HTAB_ATOMIC_STORE)
    {
        atomic_store_op(p, arg, r);
        return T(0);
    }
    else if constexpr (OP == HTAB_ATOMIC_EXCHANGE) return atomic_exchange_op(p, arg, r);
    else if constexpr (OP == HTAB_ATOMIC_COMPARE_EXCHANGE)
    {
        const T old = atomic_cas_op(p, expected, arg, r);
        ok = old == expected; // compare_exchange's result (atomic_detail.cuh:583-610)
        return old;
    }
    else if constexpr (OP == HTAB_ATOMIC_FETCH_ADD) return atomic_fetch_add_op(p, arg, r);
    else if constexpr (OP == HTAB_ATOMIC_FETCH_SUB) return atomic_fetch_sub_op(p, arg, r);
    else if constexpr (OP == HTAB_ATOMIC_FETCH_MIN) return atomic_fetch_min_op(p, arg, r);
    else if constexpr (OP == HTAB_ATOMIC_FETCH_MAX) return atomic_fetch_max_op(p, arg, r);
    else if constexpr (OP == HTAB_ATOMIC_FETCH_AND) return atomic_fetch_and_op(p, arg, r);
    else if constexpr (OP == HTAB_ATOMIC_FETCH_OR) return atomic_fetch_or_op(p, arg, r);
    else if constexpr (OP == HTAB_ATOMIC_FETCH_XOR) return atomic_fetch_xor_op(p, arg, r);
    else if constexpr (OP == HTAB_ATOMIC_FETCH_INC_MOD) return atomic_inc_wrap_op(p, arg - 1U, r);
    else return atomic_dec_wrap_op(p, arg - 1U, r);
}

__device__ inline int
highest_lane(uint64_t m)
{
    return 63 - __builtin_clzll(m);
}

constexpr int kBlock = 1024;
constexpr int kWaves = kBlock / 64;

template <typename T, int OP>
__global__ __launch_bounds__(kBlock) void
k_atomic_apply(T* __restrict__ data, int64_t count, const int64_t* __restrict__ targets, const T* __restrict__ operands,
               const T* __restrict__ expected, int64_t n, T* __restrict__ old_out, uint8_t* __restrict__ done_out)
{
    // no early exit: every thread takes part in the ballots, shuffles and barriers
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool live = i < n;
    const int64_t t = (live && targets) ? targets[i] : 0;
    const bool active = live && t >= 0 && t < count;
    const T arg = (active && OP != HTAB_ATOMIC_LOAD) ? operands[i] : T(0);
    const T exp = (active && OP == HTAB_ATOMIC_COMPARE_EXCHANGE) ? expected[i] : T(0);
    T old = T(0);
    bool ok = true;
    bool done = false;
    if constexpr (kAggregates<T, OP>)
    {
        constexpr bool kSwap = OP == HTAB_ATOMIC_EXCHANGE || OP == HTAB_ATOMIC_STORE;
        // per wavefront: 0 = no operation, 1 = every operation names element s_t0[w], 2 = mixed
        __shared__ int s_state[kWaves];
        __shared__ int64_t s_t0[kWaves];
        __shared__ T s_total[kWaves]; // the wavefront's combined operand (swap: its last operand)
        __shared__ T s_base[kWaves];  // the value the wavefront's first operation sees
        __shared__ int s_block_one;   // 1: the workgroup applied all its operations with one atomic
        const int lane = lane_id();
        const int w = static_cast<int>(threadIdx.x) >> 6;
        const uint64_t act = ballot(active);
        const int leader = act ? ctz64(act) : 0;
        const int64_t t0 = static_cast<int64_t>(shfl64(static_cast<uint64_t>(t), leader));
        const bool uniform = act != 0 && ballot(active && t != t0) == 0;
        // lane-order scan of the operands: x = combination of lanes 0..lane, ex = of lanes below lane
        T x = active ? arg : identity<T, OP>();
        T ex = identity<T, OP>();
        T total = T(0);
        const uint64_t below = act & ((uint64_t(1) << lane) - 1);
        T prev = T(0); // swap: the operand of the live lane below this one, if any
        if constexpr (kSwap)
        {
            total = act ? shfl(arg, highest_lane(act)) : T(0);
            prev = shfl(arg, below ? highest_lane(below) : lane);
        }
        else
        {
#pragma unroll
            for (int d = 1; d < 64; d <<= 1)
            {
                const T y = shfl(x, lane >= d ? lane - d : lane);
                if (lane >= d) x = combine<T, OP>(y, x);
            }
            total = shfl(x, 63);
            const T before = shfl(x, lane > 0 ? lane - 1 : 0);
            ex = lane > 0 ? before : identity<T, OP>();
        }
        if (lane == 0)
        {
            s_state[w] = act == 0 ? 0 : (uniform ? 1 : 2);
            s_t0[w] = t0;
            s_total[w] = total;
        }
        __syncthreads();
        if (threadIdx.x == 0)
        {
            // one element for the whole workgroup: one atomic applies the combination of all its
            // operations in wavefront order
            int one = 0;
            int64_t te = -1;
            for (int k = 0; k < kWaves; ++k)
            {
                if (s_state[k] == 0) continue;
                if (s_state[k] == 2 || (te >= 0 && s_t0[k] != te))
                {
                    one = -1;
                    break;
                }
                te = s_t0[k];
                one = 1;
            }
            if (one == 1)
            {
                T comb = identity<T, OP>();
                T last = T(0);
                for (int k = 0; k < kWaves; ++k)
                {
                    if (s_state[k] == 0) continue;
                    if constexpr (kSwap) last = s_total[k];
                    else comb = combine<T, OP>(comb, s_total[k]);
                }
                bool unused = true;
                const T v = apply_one<T, OP>(data + te, kSwap ? last : comb, T(0), unused);
                // the value each wavefront's first operation sees
                T run = identity<T, OP>();
                T seen = v;
                for (int k = 0; k < kWaves; ++k)
                {
                    if (s_state[k] == 0) continue;
                    if constexpr (kSwap)
                    {
                        s_base[k] = seen;
                        seen = s_total[k];
                    }
                    else
                    {
                        if constexpr (OP == HTAB_ATOMIC_FETCH_SUB)
                            s_base[k] = static_cast<T>(static_cast<std::make_unsigned_t<T>>(v) -
                                                       static_cast<std::make_unsigned_t<T>>(run));
                        else
                            s_base[k] = combine<T, OP>(v, run);
                        run = combine<T, OP>(run, s_total[k]);
                    }
                }
            }
            s_block_one = one == 1 ? 1 : 0;
        }
        __syncthreads();
        if (s_block_one || uniform)
        {
            T base;
            if (s_block_one)
                base = s_base[w];
            else
            {
                // one element for this wavefront only: one atomic for the wavefront
                T v = T(0);
                if (lane == leader)
                {
                    bool unused = true;
                    v = apply_one<T, OP>(data + t0, total, T(0), unused);
                }
                base = shfl(v, leader);
            }
            if constexpr (kSwap) old = below ? prev : base;
            else if constexpr (OP == HTAB_ATOMIC_FETCH_SUB)
                old = static_cast<T>(static_cast<std::make_unsigned_t<T>>(base) -
                                     static_cast<std::make_unsigned_t<T>>(ex));
            else old = combine<T, OP>(base, ex);
            done = active;
        }
    }
    if (!done && active) old = apply_one<T, OP>(data + t, arg, exp, ok);
    if (live)
    {
        if (old_out && OP != HTAB_ATOMIC_STORE) old_out[i] = old;
        if (done_out) done_out[i] = (active && ok) ? 1 : 0;
    }
}

// the operations the reference defines for T (atomic.cuh:295-377): bitwise ones for the integer
// types only, the wrapping increment / decrement for unsigned int only
template <typename T, int OP>
constexpr bool kDefined = (OP <= HTAB_ATOMIC_FETCH_MAX) ||
                          ((OP == HTAB_ATOMIC_FETCH_AND || OP == HTAB_ATOMIC_FETCH_OR || OP == HTAB_ATOMIC_FETCH_XOR) &&
                           !std::is_same_v<T, float>) ||
                          ((OP == HTAB_ATOMIC_FETCH_INC_MOD || OP == HTAB_ATOMIC_FETCH_DEC_MOD) &&
                           std::is_same_v<T, unsigned int>);

template <typename T, int OP>
int
launch_op(void* data, int64_t count, const int64_t* targets, const void* operands, const void* expected, int64_t n,
          void* old_out, uint8_t* done_out, hipStream_t s)
{
    if constexpr (!kDefined<T, OP>) return HTAB_E_ARG;
    else
    {
        const int64_t grid = (n + kBlock - 1) / kBlock;
        if (grid > 0x7FFFFFFF) return HTAB_E_LIMIT;
        hipLaunchKernelGGL((k_atomic_apply<T, OP>), dim3(static_cast<uint32_t>(grid)), dim3(kBlock), 0, s,
                           static_cast<T*>(data), count, targets, static_cast<const T*>(operands),
                           static_cast<const T*>(expected), n, static_cast<T*>(old_out), done_out);
        return hipGetLastError() == hipSuccess ? HTAB_OK : HTAB_E_HIP;
    }
}

template <typename T>
int
dispatch_op(int op, void* data, int64_t count, const int64_t* targets, const void* operands, const void* expected,
            int64_t n, void* old_out, uint8_t* done_out, hipStream_t s)
{
#define HTAB_ATOMIC_CASE(OPC)                                                                                         \
    case OPC:                                                                                                        \
        return launch_op<T, OPC>(data, count, targets, operands, expected, n, old_out, done_out, s);
    switch (op)
    {
        HTAB_ATOMIC_CASE(HTAB_ATOMIC_LOAD)
        HTAB_ATOMIC_CASE(HTAB_ATOMIC_STORE)
        HTAB_ATOMIC_CASE(HTAB_ATOMIC_EXCHANGE)
        HTAB_ATOMIC_CASE(HTAB_ATOMIC_COMPARE_EXCHANGE)
        HTAB_ATOMIC_CASE(HTAB_ATOMIC_FETCH_ADD)
        HTAB_ATOMIC_CASE(HTAB_ATOMIC_FETCH_SUB)
        HTAB_ATOMIC_CASE(HTAB_ATOMIC_FETCH_MIN)
        HTAB_ATOMIC_CASE(HTAB_ATOMIC_FETCH_MAX)
        HTAB_ATOMIC_CASE(HTAB_ATOMIC_FETCH_AND)
        HTAB_ATOMIC_CASE(HTAB_ATOMIC_FETCH_OR)
        HTAB_ATOMIC_CASE(HTAB_ATOMIC_FETCH_XOR)
        HTAB_ATOMIC_CASE(HTAB_ATOMIC_FETCH_INC_MOD)
        HTAB_ATOMIC_CASE(HTAB_ATOMIC_FETCH_DEC_MOD)
    default:
        return HTAB_E_ARG;
    }
#undef HTAB_ATOMIC_CASE
}

size_t
type_bytes(int type)
{
    switch (type)
    {
    case HTAB_ATOMIC_I32:
    case HTAB_ATOMIC_U32:
    case HTAB_ATOMIC_F32:
        return 4;
    case HTAB_ATOMIC_U64:
        return 8;
    default:
        return 0;
    }
}

} // namespace

struct htab_atomic
{
    void* data = nullptr;
    int type = 0;
    int64_t count = 0;
};

extern "C" {

int
htab_atomic_create(int type, int64_t count, htab_atomic** out)
{
    if (!out || type_bytes(type) == 0 || count <= 0 || count > (int64_t(1) << 36)) return HTAB_E_ARG;
    *out = nullptr;
    htab_atomic* a = new (std::nothrow) htab_atomic();
    if (!a) return HTAB_E_OOM;
    const size_t bytes = static_cast<size_t>(count) * type_bytes(type);
    if (hipMalloc(&a->data, bytes) != hipSuccess)
    {
        delete a;
        return HTAB_E_OOM;
    }
    // createDeviceObject fills the value with 0 (atomic_detail.cuh:143-151)
    if (hipMemset(a->data, 0, bytes) != hipSuccess)
    {
        (void)hipFree(a->data);
        delete a;
        return HTAB_E_HIP;
    }
    a->type = type;
    a->count = count;
    *out = a;
    return HTAB_OK;
}

int
htab_atomic_destroy(htab_atomic* a)
{
    if (!a) return HTAB_E_ARG;
    const hipError_t e = hipFree(a->data);
    delete a;
    return e == hipSuccess ? HTAB_OK : HTAB_E_HIP;
}

int
htab_atomic_data(const htab_atomic* a, void** out)
{
    if (!a || !out) return HTAB_E_ARG;
    *out = a->data;
    return HTAB_OK;
}

int
htab_atomic_ref_apply(int type, void* data, int64_t count, int op, const int64_t* targets, const void* operands,
                      const void* expected, int64_t n, void* old_out, uint8_t* done_out, void* stream)
{
    if (type_bytes(type) == 0 || !data || count <= 0 || n < 0) return HTAB_E_ARG;
    // device atomics and loads need naturally aligned values, operands and targets
    const uintptr_t w = type_bytes(type);
    for (const void* p : { static_cast<const void*>(data), operands, expected, static_cast<const void*>(old_out) })
        if (p && reinterpret_cast<uintptr_t>(p) % w != 0) return HTAB_E_ARG;
    if (targets && reinterpret_cast<uintptr_t>(targets) % 8 != 0) return HTAB_E_ARG;
    if (op < HTAB_ATOMIC_LOAD || op > HTAB_ATOMIC_FETCH_DEC_MOD) return HTAB_E_ARG;
    if (n > 0 && op != HTAB_ATOMIC_LOAD && !operands) return HTAB_E_ARG;
    if (n > 0 && op == HTAB_ATOMIC_COMPARE_EXCHANGE && !expected) return HTAB_E_ARG;
    if (n == 0) return HTAB_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    switch (type)
    {
    case HTAB_ATOMIC_I32:
        return dispatch_op<int>(op, data, count, targets, operands, expected, n, old_out, done_out, s);
    case HTAB_ATOMIC_U32:
        return dispatch_op<unsigned int>(op, data, count, targets, operands, expected, n, old_out, done_out, s);
    case HTAB_ATOMIC_U64:
        return dispatch_op<unsigned long long>(op, data, count, targets, operands, expected, n, old_out, done_out, s);
    default:
        return dispatch_op<float>(op, data, count, targets, operands, expected, n, old_out, done_out, s);
    }
}

int
htab_atomic_apply(htab_atomic* a, int op, const int64_t* targets, const void* operands, const void* expected,
                  int64_t n, void* old_out, uint8_t* done_out, void* stream)
{
    if (!a) return HTAB_E_ARG;
    return htab_atomic_ref_apply(a->type, a->data, a->count, op, targets, operands, expected, n, old_out, done_out,
                                 stream);
}

int
htab_atomic_load(const htab_atomic* a, int64_t index, void* value_out, void* stream)
{
    if (!a || !value_out || index < 0 || index >= a->count) return HTAB_E_ARG;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const size_t w = type_bytes(a->type);
    if (hipMemcpyAsync(value_out, static_cast<char*>(a->data) + static_cast<size_t>(index) * w, w, hipMemcpyDefault,
                       s) != hipSuccess)
        return HTAB_E_HIP;
    return hipStreamSynchronize(s) == hipSuccess ? HTAB_OK : HTAB_E_HIP;
}

int
htab_atomic_store(htab_atomic* a, int64_t index, const void* value, void* stream)
{
    if (!a || !value || index < 0 || index >= a->count) return HTAB_E_ARG;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const size_t w = type_bytes(a->type);
    if (hipMemcpyAsync(static_cast<char*>(a->data) + static_cast<size_t>(index) * w, value, w, hipMemcpyDefault, s) !=
        hipSuccess)
        return HTAB_E_HIP;
    return hipStreamSynchronize(s) == hipSuccess ? HTAB_OK : HTAB_E_HIP;
}

} // extern "C"
