// CPU restatement of stdgpu's hashed-container engine (OpenMP backend), used ONLY as
// test infrastructure: the parity checker for the HIP library and the timed CPU
// baseline.  Nothing in the product links or calls this file.
//
// What it restates (reference file:line, paths relative to the stdgpu source tree):
//   * sizing            src/stdgpu/impl/unordered_base_detail.cuh:34-54, 1254-1308
//                       src/stdgpu/impl/bit_detail.h:34-55, 92-110
//   * Fibonacci hashing src/stdgpu/impl/unordered_base_detail.cuh:56-85
//   * integer hash      src/stdgpu/impl/functional_detail.h:60-80 (identity)
//   * layout            src/stdgpu/impl/unordered_base.cuh:526-554
//   * find              src/stdgpu/impl/unordered_base_detail.cuh:653-682
//   * try_insert/insert src/stdgpu/impl/unordered_base_detail.cuh:708-809, 967-994
//   * try_erase/erase   src/stdgpu/impl/unordered_base_detail.cuh:811-956, 1019-1043
//   * excess free-list  src/stdgpu/impl/vector_detail.cuh:181-302, 361-368
//   * lock bits         src/stdgpu/impl/mutex_detail.cuh:31-46, 153-191
//   * occupancy bits    src/stdgpu/impl/bitset_detail.cuh:47-84
//   * deque free-list   src/stdgpu/impl/deque_detail.cuh:202-260, 390-512
//   * valid()           src/stdgpu/impl/unordered_base_detail.cuh:210-468, 1192-1218
//   * atomics cost      src/stdgpu/openmp/impl/atomic_detail.h:33-219: load/store/exchange/CAS/inc_mod
//                       are one global `omp critical`; fetch_add/sub/and/or are `omp atomic capture`
//   * bulk launch       src/stdgpu/impl/algorithm_detail.h:52-63 (for_each_index -> omp parallel for)
//
// Parity pins (checked by tests/test_oracle.py): closed-form sizing and bucket() values
// recorded from the compiled reference, the reference tests' pinned colliding keys,
// brute-force set semantics, the reference's valid() invariants, and the examples'
// closed-form sums.
#include <omp.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <type_traits>

namespace oracle
{

using index_t = std::int32_t;

// ---------------------------------------------------------------- atomics (OpenMP backend cost model)
// Every atomic_ref operation of the reference defaults to memory_order_seq_cst, which puts a
// backend fence (`omp flush`) before and after the backend call (atomic_detail.cuh:34-97 and
// :436-456, 490-505, 540-600; openmp/impl/atomic_detail.h:33-36).  The fences order nothing that
// the critical sections do not already order here, but they are part of what the reference's
// OpenMP path costs, and this restatement is also the timed CPU baseline.
static inline void
seq_cst_fence()
{
#pragma omp flush
}

template <typename T>
static inline T
a_load(const T* p)
{
    T v;
    seq_cst_fence();
#pragma omp critical
    v = *p;
    seq_cst_fence();
    return v;
}

template <typename T>
static inline void
a_store(T* p, T v)
{
    seq_cst_fence();
#pragma omp critical
    *p = v;
    seq_cst_fence();
}

template <typename T>
static inline T
a_fetch_add(T* p, T arg)
{
    T old;
    seq_cst_fence();
#pragma omp atomic capture
    {
        old = *p;
        *p += arg;
    }
    seq_cst_fence();
    return old;
}

template <typename T>
static inline T
a_fetch_sub(T* p, T arg)
{
    T old;
    seq_cst_fence();
#pragma omp atomic capture
    {
        old = *p;
        *p -= arg;
    }
    seq_cst_fence();
    return old;
}

template <typename T>
static inline T
a_fetch_or(T* p, T arg)
{
    T old;
    seq_cst_fence();
#pragma omp atomic capture
    {
        old = *p;
        *p |= arg;
    }
    seq_cst_fence();
    return old;
}

template <typename T>
static inline T
a_fetch_and(T* p, T arg)
{
    T old;
    seq_cst_fence();
#pragma omp atomic capture
    {
        old = *p;
        *p &= arg;
    }
    seq_cst_fence();
    return old;
}

// atomicInc-style wrap: stored value goes 0,1,..,mod-1,0,...  (caller passes capacity, backend gets capacity-1)
static inline unsigned int
a_fetch_inc_mod(unsigned int* p, unsigned int capacity)
{
    unsigned int old;
    const unsigned int arg = capacity - 1;
    seq_cst_fence();
#pragma omp critical
    {
        old = *p;
        *p = (old >= arg) ? 0U : old + 1U;
    }
    seq_cst_fence();
    return old;
}

// atomicDec-style wrap (openmp/impl/atomic_detail.h:205-216): stored value goes mod-1,...,1,0,mod-1,...
// (caller passes capacity, backend gets capacity-1, atomic_detail.cuh:693-699)
static inline unsigned int
a_fetch_dec_mod(unsigned int* p, unsigned int capacity)
{
    unsigned int old;
    const unsigned int arg = capacity - 1;
    seq_cst_fence();
#pragma omp critical
    {
        old = *p;
        *p = (old == 0U || old > arg) ? arg : old - 1U;
    }
    seq_cst_fence();
    return old;
}

// ---------------------------------------------------------------- bit helpers (bit_detail.h)
static inline std::uint64_t
bit_ceil(std::uint64_t x)
{
    if (x <= 1) return 1;
    std::uint64_t r = 1;
    while (r < x) r <<= 1;
    return r;
}

static inline int
bit_width(std::uint64_t x)
{
    int w = 0;
    while (x) { ++w; x >>= 1; }
    return w;
}

// unordered_base_detail.cuh:56-85
static inline index_t
fibonacci_hashing(std::uint64_t hash, index_t bucket_count)
{
    const int w = bit_width(static_cast<std::uint64_t>(bucket_count)) - 1;
    if (w <= 0) return 0;
    const unsigned dropped = static_cast<unsigned>(64 - w);
    const std::uint64_t improved = hash ^ (hash >> dropped);
    const std::uint64_t golden = 11400714819323198485ULL; // floor(2^64 / phi)
    return static_cast<index_t>((golden * improved) >> dropped);
}

// unordered_base_detail.cuh:34-48 (long double arithmetic, as in the reference)
static inline index_t
expected_collisions(index_t bucket_count, index_t capacity)
{
    const long double k = static_cast<long double>(bucket_count);
    const long double n = static_cast<long double>(capacity);
    return static_cast<index_t>(std::ceil(n * (1.0L - std::pow(1.0L - (1.0L / k), n - 1.0L))));
}

// unordered_base_detail.cuh:1263-1273 (note the float conversion of capacity)
static inline void
sizing(index_t capacity, index_t* bucket_count, index_t* excess_count)
{
    const float maxlf = 1.0F;
    const std::uint64_t want = static_cast<std::uint64_t>(std::ceil(static_cast<float>(capacity) / maxlf));
    *bucket_count = static_cast<index_t>(bit_ceil(want));
    *excess_count = expected_collisions(*bucket_count, capacity) + 1;
}

// ---------------------------------------------------------------- atomic bit array (bitset_detail.cuh:47-84)
struct bits
{
    unsigned int* w = nullptr;
    index_t n = 0;

    void init(index_t count, bool value)
    {
        n = count;
        const index_t nw = (count + 31) / 32;
        w = static_cast<unsigned int*>(std::malloc(sizeof(unsigned int) * static_cast<std::size_t>(nw > 0 ? nw : 1)));
        std::memset(w, value ? 0xFF : 0, sizeof(unsigned int) * static_cast<std::size_t>(nw > 0 ? nw : 1));
    }
    void release() { std::free(w); w = nullptr; }
    bool test(index_t i) const { return (a_load(&w[i >> 5]) >> (i & 31)) & 1U; }
    // returns the previous bit, like bitset::reference::operator=(true)
    bool set(index_t i) { const unsigned m = 1U << (i & 31); return (a_fetch_or(&w[i >> 5], m) & m) != 0; }
    bool reset(index_t i) { const unsigned m = 1U << (i & 31); return (a_fetch_and(&w[i >> 5], ~m) & m) != 0; }
    // mutex_array semantics (mutex_detail.cuh:31-46)
    bool try_lock(index_t i) { return !set(i); }
    void unlock(index_t i) { reset(i); }
    index_t count() const
    {
        index_t c = 0;
        for (index_t i = 0; i < n; ++i) c += (w[i >> 5] >> (i & 31)) & 1U;
        return c;
    }
};

// all-or-nothing two-lock acquire, -1 on success (mutex_detail.cuh:153-191)
static inline int
try_lock_pair(bits& locks, index_t a, index_t b)
{
    if (!locks.try_lock(a)) return 0;
    if (!locks.try_lock(b))
    {
        locks.unlock(a);
        return 1;
    }
    return -1;
}

// ---------------------------------------------------------------- vector<index_t> used as a LIFO stack
struct index_stack
{
    index_t* data = nullptr;
    index_t cap = 0;
    int size_ = 0;
    bits occ, locks;

    void init(index_t capacity)
    {
        cap = capacity;
        data = static_cast<index_t*>(std::malloc(sizeof(index_t) * static_cast<std::size_t>(capacity > 0 ? capacity : 1)));
        occ.init(capacity, false);
        locks.init(capacity, false);
    }
    void release() { std::free(data); occ.release(); locks.release(); }
    // vector_detail.cuh:454-512 (clamped size)
    index_t size() const
    {
        const int s = a_load(&size_);
        if (s < 0) return 0;
        if (s > cap) return cap;
        return s;
    }
    bool empty() const { return size() == 0; }
    bool full() const { return size() == cap; }
    // vector_clear_iota (vector_detail.cuh:361-368)
    void clear_iota(index_t first)
    {
        for (index_t i = 0; i < cap; ++i) data[i] = first + i;
        const index_t nw = (cap + 31) / 32;
        for (index_t i = 0; i < nw; ++i) occ.w[i] = 0xFFFFFFFFU;
        a_store(&size_, static_cast<int>(cap));
    }
    // vector_detail.cuh:242-302
    bool pop_back(index_t* out)
    {
        if (empty()) return false;
        const index_t pos = a_fetch_sub(&size_, 1) - 1;
        if (pos < 0 || pos >= cap) return false;
        bool popped = false;
        while (!popped)
        {
            if (locks.try_lock(pos))
            {
                if (occ.test(pos))
                {
                    occ.reset(pos);
                    *out = data[pos];
                    popped = true;
                }
                locks.unlock(pos);
            }
        }
        return true;
    }
    // vector_detail.cuh:181-240
    bool push_back(index_t v)
    {
        if (full()) return false;
        const index_t pos = a_fetch_add(&size_, 1);
        if (pos < 0 || pos >= cap) return false;
        bool pushed = false;
        while (!pushed)
        {
            if (locks.try_lock(pos))
            {
                if (!occ.test(pos))
                {
                    data[pos] = v;
                    occ.set(pos);
                    pushed = true;
                }
                locks.unlock(pos);
            }
        }
        return true;
    }
};

// ---------------------------------------------------------------- value types
template <typename K, typename M>
struct kv
{
    K first;
    M second;
};

template <typename K, typename M>
struct traits
{
    using value_type = kv<K, M>;
    static K key(const value_type& v) { return v.first; }
    static void make(value_type* dst, K k, const void* vals, std::int64_t i)
    {
        dst->first = k;
        dst->second = vals ? static_cast<const M*>(vals)[i] : M{};
    }
    static void out(void* vals, std::int64_t i, const value_type& v) { static_cast<M*>(vals)[i] = v.second; }
    static constexpr bool is_map = true;
};

template <typename K>
struct traits<K, void>
{
    using value_type = K;
    static K key(const value_type& v) { return v; }
    static void make(value_type* dst, K k, const void*, std::int64_t) { *dst = k; }
    static void out(void*, std::int64_t, const value_type&) {}
    static constexpr bool is_map = false;
};

enum status
{
    success,
    failed_collision,
    failed_no_action_required
};

struct table_base
{
    virtual ~table_base() = default;
    virtual void insert(const void* keys, const void* vals, std::int64_t n, std::uint8_t* ok, int threads) = 0;
    virtual void find(const void* keys, std::int64_t n, std::uint8_t* hits, void* vals, int threads) const = 0;
    virtual void erase(const void* keys, std::int64_t n, std::uint8_t* ok, int threads) = 0;
    virtual std::int64_t size() const = 0;
    virtual std::int64_t contents(void* keys, void* vals) const = 0;
    virtual int valid() const = 0;
    virtual void clear() = 0;
    virtual std::int64_t excess_free() const = 0;
    virtual int corrupt(int what) = 0;
    index_t bucket_count_ = 0;
    index_t total_ = 0;
    virtual index_t bucket(std::uint64_t key) const = 0;
};

template <typename K, typename M>
struct table final : table_base
{
    using T = traits<K, M>;
    using value_type = typename T::value_type;

    value_type* values = nullptr;
    index_t* offsets = nullptr;
    bits occupied, locks;
    index_stack excess;
    int occupied_count = 0;

    explicit table(index_t capacity)
    {
        index_t bc = 0, ex = 0;
        sizing(capacity, &bc, &ex);
        bucket_count_ = bc;
        total_ = bc + ex;
        values = static_cast<value_type*>(std::calloc(static_cast<std::size_t>(total_), sizeof(value_type)));
        offsets = static_cast<index_t*>(std::calloc(static_cast<std::size_t>(total_), sizeof(index_t)));
        occupied.init(total_, false);
        locks.init(total_, false);
        excess.init(ex);
        excess.clear_iota(bc);
    }
    ~table() override
    {
        std::free(values);
        std::free(offsets);
        occupied.release();
        locks.release();
        excess.release();
    }

    // hash<integer> is the identity cast to size_t (functional_detail.h:60-80)
    index_t bucket_of(K key) const { return fibonacci_hashing(static_cast<std::uint64_t>(key), bucket_count_); }
    index_t bucket(std::uint64_t key) const override { return bucket_of(static_cast<K>(key)); }

    index_t size_now() const { return a_load(&occupied_count); }
    bool full() const { return size_now() == total_; }

    // find_impl (unordered_base_detail.cuh:653-682); returns total_ for end()
    index_t find_pos(K key) const
    {
        index_t i = bucket_of(key);
        if (occupied.test(i) && T::key(values[i]) == key) return i;
        while (offsets[i] != 0)
        {
            i += offsets[i];
            if (occupied.test(i) && T::key(values[i]) == key) return i;
        }
        return total_;
    }
    bool contains(K key) const { return find_pos(key) != total_; }

    index_t list_end(index_t start) const
    {
        index_t e = start;
        while (offsets[e] != 0) e += offsets[e];
        return e;
    }

    index_t previous_position(index_t pos, index_t start) const
    {
        index_t prev = start, i = start;
        while (offsets[i] != 0)
        {
            i += offsets[i];
            if (i == pos) break;
            prev = i;
        }
        return prev;
    }

    // unordered_base_detail.cuh:708-809
    status try_insert(const value_type& v)
    {
        const K key = T::key(v);
        if (contains(key)) return failed_no_action_required;
        const index_t b = bucket_of(key);
        if (!occupied.test(b))
        {
            if (locks.try_lock(b))
            {
                status st = failed_collision;
                if (!contains(key) && !occupied.test(b))
                {
                    values[b] = v;
                    a_fetch_add(&occupied_count, 1);
                    occupied.set(b);
                    st = success;
                }
                locks.unlock(b);
                return st;
            }
            return failed_collision;
        }
        const index_t end = list_end(b);
        if (locks.try_lock(end))
        {
            status st = failed_collision;
            const index_t end2 = list_end(b);
            if (!contains(key) && end == end2)
            {
                index_t slot = 0;
                if (excess.pop_back(&slot))
                {
                    values[slot] = v;
                    offsets[slot] = 0;
                    a_fetch_add(&occupied_count, 1);
                    occupied.set(slot);
                    offsets[end] = slot - end; // link after the entry is fully initialised
                    st = success;
                }
            }
            locks.unlock(end);
            return st;
        }
        return failed_collision;
    }

    // unordered_base_detail.cuh:967-994
    bool insert_one(const value_type& v)
    {
        status st = failed_collision;
        while (st == failed_collision && !full() && !excess.empty()) st = try_insert(v);
        return st == success;
    }

    // unordered_base_detail.cuh:811-913
    status try_erase(K key)
    {
        const index_t pos = find_pos(key);
        if (pos == total_) return failed_no_action_required;
        const index_t b = bucket_of(key);
        if (pos == b)
        {
            if (locks.try_lock(pos))
            {
                status st = failed_collision;
                if (find_pos(key) == pos)
                {
                    occupied.reset(pos);
                    a_fetch_sub(&occupied_count, 1);
                    st = success;
                }
                locks.unlock(pos);
                return st;
            }
            return failed_collision;
        }
        const index_t prev = previous_position(pos, b);
        if (try_lock_pair(locks, pos, prev) == -1)
        {
            status st = failed_collision;
            if (find_pos(key) == pos && previous_position(pos, b) == prev)
            {
                if (offsets[pos] != 0)
                    offsets[prev] += offsets[pos];
                else
                    offsets[prev] = 0;
                occupied.reset(pos);
                a_fetch_sub(&occupied_count, 1);
                excess.push_back(pos);
                st = success;
            }
            locks.unlock(pos);
            locks.unlock(prev);
            return st;
        }
        return failed_collision;
    }

    bool erase_one(K key)
    {
        status st = failed_collision;
        while (st == failed_collision) st = try_erase(key);
        return st == success;
    }

    void insert(const void* keys, const void* vals, std::int64_t n, std::uint8_t* ok, int threads) override
    {
        const K* k = static_cast<const K*>(keys);
#pragma omp parallel for num_threads(threads) schedule(static)
        for (std::int64_t i = 0; i < n; ++i)
        {
            value_type v;
            T::make(&v, k[i], vals, i);
            const bool r = insert_one(v);
            if (ok) ok[i] = r ? 1 : 0;
        }
    }

    void find(const void* keys, std::int64_t n, std::uint8_t* hits, void* vals, int threads) const override
    {
        const K* k = static_cast<const K*>(keys);
#pragma omp parallel for num_threads(threads) schedule(static)
        for (std::int64_t i = 0; i < n; ++i)
        {
            const index_t p = find_pos(k[i]);
            const bool h = p != total_;
            if (hits) hits[i] = h ? 1 : 0;
            if (vals && h) T::out(vals, i, values[p]);
        }
    }

    void erase(const void* keys, std::int64_t n, std::uint8_t* ok, int threads) override
    {
        const K* k = static_cast<const K*>(keys);
#pragma omp parallel for num_threads(threads) schedule(static)
        for (std::int64_t i = 0; i < n; ++i)
        {
            const bool r = erase_one(k[i]);
            if (ok) ok[i] = r ? 1 : 0;
        }
    }

    std::int64_t size() const override { return size_now(); }

    std::int64_t contents(void* keys, void* vals) const override
    {
        std::int64_t c = 0;
        for (index_t i = 0; i < total_; ++i)
        {
            if (!occupied.test(i)) continue;
            if (keys) static_cast<K*>(keys)[c] = T::key(values[i]);
            if (vals) T::out(vals, c, values[i]);
            ++c;
        }
        return c;
    }

    // valid() (unordered_base_detail.cuh:1192-1218): the reference's seven checks in its order,
    // each stopping the rest (its && chain), so a looping chain never reaches the finds below.
    int valid() const override
    {
        if (total_ <= 0) return 1;
        // offset_range_valid (:211-256): every linked entry i + offsets[i] lies inside the table
        for (index_t i = 0; i < total_; ++i)
        {
            const std::int64_t linked = static_cast<std::int64_t>(i) + offsets[i];
            if (linked < 0 || linked >= total_) return 0;
        }
        // loop_free (:259-351): walk the chain of every bucket, counting visits of each position,
        // the bucket head included; a walk stops at a position seen twice; all counts must be <= 1
        int* visits = static_cast<int*>(std::calloc(static_cast<std::size_t>(total_), sizeof(int)));
        int ok = 1;
        for (index_t b = 0; b < bucket_count_; ++b)
        {
            index_t i = b;
            ++visits[i];
            while (offsets[i] != 0)
            {
                i += offsets[i];
                if (++visits[i] > 1) break;
            }
        }
        for (index_t i = 0; i < total_ && ok; ++i)
            if (visits[i] > 1) ok = 0;
        std::free(visits);
        if (!ok) return 0;
        // values_reachable (:353-399): every occupied value is found by contains()
        for (index_t i = 0; i < total_; ++i)
            if (occupied.test(i) && !contains(T::key(values[i]))) return 0;
        // unique (:401-452): find() of every occupied value returns its own position
        for (index_t i = 0; i < total_; ++i)
            if (occupied.test(i) && find_pos(T::key(values[i])) != i) return 0;
        // occupied_count_valid (:454-468): size() equals the number of set occupancy bits
        if (occupied.count() != a_load(&occupied_count)) return 0;
        // _locks.valid() (mutex_detail.cuh:136-152): no lock bit set
        if (locks.count() != 0) return 0;
        // _excess_list_positions.valid() (vector_detail.cuh:614-636, 776-796): size inside
        // [0, capacity], size() equal to the occupancy count, no lock bit set
        if (excess.cap != 0)
        {
            const int raw = a_load(&excess.size_);
            if (raw < 0 || raw > excess.cap) return 0;
            if (excess.occ.count() != excess.size()) return 0;
            if (excess.locks.count() != 0) return 0;
        }
        return 1;
    }

    // Test-only: breaks one invariant of a filled table so that a test can require valid() == 0
    // for it.  Returns 1 when the table had the structure the case needs (a chain with an excess
    // entry, a free excess position), else 0 and leaves the table unchanged.
    int corrupt(int what) override
    {
        // a bucket whose chain holds at least one excess entry: head b, first linked entry e
        index_t b = -1;
        for (index_t i = 0; i < bucket_count_; ++i)
            if (offsets[i] != 0 && occupied.test(i)) { b = i; break; }
        switch (what)
        {
            case 0: // loop: the chain's last entry links back to its bucket head
            {
                if (b < 0) return 0;
                const index_t end = list_end(b);
                offsets[end] = b - end;
                return 1;
            }
            case 1: // an offset pointing past the last position
            {
                if (b < 0) return 0;
                const index_t end = list_end(b);
                offsets[end] = total_ - end;
                return 1;
            }
            case 2: // an occupied entry that no chain reaches (size kept equal to the bit count)
            {
                if (excess.size() == 0) return 0;
                const index_t pos = excess.data[excess.size() - 1]; // a free excess position
                K fresh = static_cast<K>(1);
                while (contains(fresh) || bucket_of(fresh) == pos) fresh = static_cast<K>(fresh + 1);
                T::make(&values[pos], fresh, nullptr, -1);
                occupied.set(pos);
                a_fetch_add(&occupied_count, 1);
                return 1;
            }
            case 3: // a duplicate: the first excess entry of a chain gets its head's key
            {
                if (b < 0) return 0;
                const index_t e = b + offsets[b];
                set_key(&values[e], T::key(values[b]));
                return 1;
            }
            case 4: // size() one above the number of occupied entries
                a_fetch_add(&occupied_count, 1);
                return 1;
            case 5: // a lock left held
                locks.set(total_ / 2);
                return 1;
            case 6: // the excess list's size one above its occupancy count
                if (excess.cap == 0 || excess.full()) return 0;
                a_fetch_add(&excess.size_, 1);
                return 1;
            default: return 0;
        }
    }

    static void set_key(value_type* v, K k)
    {
        if constexpr (T::is_map)
            v->first = k;
        else
            *v = k;
    }

    // unordered_base_detail.cuh:1220-1252
    void clear() override
    {
        std::memset(offsets, 0, sizeof(index_t) * static_cast<std::size_t>(total_));
        std::memset(occupied.w, 0, sizeof(unsigned int) * static_cast<std::size_t>((total_ + 31) / 32));
        a_store(&occupied_count, 0);
        std::memset(excess.occ.w, 0, sizeof(unsigned int) * static_cast<std::size_t>((excess.cap + 31) / 32));
        excess.clear_iota(bucket_count_);
    }

    std::int64_t excess_free() const override { return excess.size(); }
};

// ---------------------------------------------------------------- deque<uint32_t> free-list (deque_detail.cuh)
struct u32_deque
{
    unsigned int* data = nullptr;
    index_t cap = 0;
    int size_ = 0;
    unsigned int begin_ = 0, end_ = 0;
    bits occ, locks;

    explicit u32_deque(index_t capacity)
    {
        cap = capacity;
        data = static_cast<unsigned int*>(std::calloc(static_cast<std::size_t>(capacity), sizeof(unsigned int)));
        occ.init(capacity, false);
        locks.init(capacity, false);
    }
    ~u32_deque()
    {
        std::free(data);
        occ.release();
        locks.release();
    }
    index_t size() const
    {
        const int s = a_load(&size_);
        if (s < 0) return 0;
        if (s > cap) return cap;
        return s;
    }
    // deque_detail.cuh:202-260
    bool push_back(unsigned int v)
    {
        if (size() == cap) return false;
        const index_t cur = a_fetch_add(&size_, 1);
        if (cur >= cap) return false;
        const index_t pos = static_cast<index_t>(a_fetch_inc_mod(&end_, static_cast<unsigned int>(cap)));
        bool pushed = false;
        while (!pushed)
        {
            if (locks.try_lock(pos))
            {
                if (!occ.test(pos))
                {
                    data[pos] = v;
                    occ.set(pos);
                    pushed = true;
                }
                locks.unlock(pos);
            }
        }
        return true;
    }
    // deque_detail.cuh:262-320: the position is the stored end minus one, modulo the capacity
    bool pop_back(unsigned int* out)
    {
        if (size() == 0) return false;
        const index_t cur = a_fetch_sub(&size_, 1);
        if (cur <= 0) return false;
        index_t pos = static_cast<index_t>(a_fetch_dec_mod(&end_, static_cast<unsigned int>(cap)));
        pos = (pos == 0) ? cap - 1 : pos - 1;
        bool popped = false;
        while (!popped)
        {
            if (locks.try_lock(pos))
            {
                if (occ.test(pos))
                {
                    occ.reset(pos);
                    *out = data[pos];
                    popped = true;
                }
                locks.unlock(pos);
            }
        }
        return true;
    }
    // deque_detail.cuh:330-388: the position is the stored begin minus one, modulo the capacity
    bool push_front(unsigned int v)
    {
        if (size() == cap) return false;
        const index_t cur = a_fetch_add(&size_, 1);
        if (cur >= cap) return false;
        index_t pos = static_cast<index_t>(a_fetch_dec_mod(&begin_, static_cast<unsigned int>(cap)));
        pos = (pos == 0) ? cap - 1 : pos - 1;
        bool pushed = false;
        while (!pushed)
        {
            if (locks.try_lock(pos))
            {
                if (!occ.test(pos))
                {
                    data[pos] = v;
                    occ.set(pos);
                    pushed = true;
                }
                locks.unlock(pos);
            }
        }
        return true;
    }
    // deque_detail.cuh:390-448
    bool pop_front(unsigned int* out)
    {
        if (size() == 0) return false;
        const index_t cur = a_fetch_sub(&size_, 1);
        if (cur <= 0) return false;
        const index_t pos = static_cast<index_t>(a_fetch_inc_mod(&begin_, static_cast<unsigned int>(cap)));
        bool popped = false;
        while (!popped)
        {
            if (locks.try_lock(pos))
            {
                if (occ.test(pos))
                {
                    occ.reset(pos);
                    *out = data[pos];
                    popped = true;
                }
                locks.unlock(pos);
            }
        }
        return true;
    }
};

} // namespace oracle

using namespace oracle;

extern "C" {

// kind: 0 set<u32>, 1 set<u64>, 2 map<u64,u32>, 3 map<u32,u32>, 4 map<u64,u64>
void*
orc_create(int kind, std::int64_t capacity)
{
    if (capacity <= 0 || capacity > 0x7FFFFFFF) return nullptr;
    const index_t c = static_cast<index_t>(capacity);
    switch (kind)
    {
        case 0: return new table<std::uint32_t, void>(c);
        case 1: return new table<std::uint64_t, void>(c);
        case 2: return new table<std::uint64_t, std::uint32_t>(c);
        case 3: return new table<std::uint32_t, std::uint32_t>(c);
        case 4: return new table<std::uint64_t, std::uint64_t>(c);
        default: return nullptr;
    }
}

void orc_destroy(void* t) { delete static_cast<table_base*>(t); }

void
orc_insert(void* t, const void* keys, const void* vals, std::int64_t n, std::uint8_t* ok, int threads)
{
    static_cast<table_base*>(t)->insert(keys, vals, n, ok, threads > 0 ? threads : 1);
}

void
orc_find(void* t, const void* keys, std::int64_t n, std::uint8_t* hits, void* vals, int threads)
{
    static_cast<table_base*>(t)->find(keys, n, hits, vals, threads > 0 ? threads : 1);
}

void
orc_erase(void* t, const void* keys, std::int64_t n, std::uint8_t* ok, int threads)
{
    static_cast<table_base*>(t)->erase(keys, n, ok, threads > 0 ? threads : 1);
}

std::int64_t orc_size(void* t) { return static_cast<table_base*>(t)->size(); }
std::int64_t orc_contents(void* t, void* keys, void* vals) { return static_cast<table_base*>(t)->contents(keys, vals); }
int orc_valid(void* t) { return static_cast<table_base*>(t)->valid(); }
void orc_clear(void* t) { static_cast<table_base*>(t)->clear(); }
int orc_corrupt(void* t, int what) { return static_cast<table_base*>(t)->corrupt(what); }
std::int64_t orc_bucket_count(void* t) { return static_cast<table_base*>(t)->bucket_count_; }
std::int64_t orc_max_size(void* t) { return static_cast<table_base*>(t)->total_; }
std::int64_t orc_excess_free(void* t) { return static_cast<table_base*>(t)->excess_free(); }
std::int64_t orc_bucket(void* t, std::uint64_t key) { return static_cast<table_base*>(t)->bucket(key); }

std::int64_t
orc_fibonacci_hashing(std::uint64_t hash, std::int64_t bucket_count)
{
    return fibonacci_hashing(hash, static_cast<index_t>(bucket_count));
}

void
orc_sizing(std::int64_t capacity, std::int64_t* bucket_count, std::int64_t* excess_count)
{
    index_t bc = 0, ex = 0;
    sizing(static_cast<index_t>(capacity), &bc, &ex);
    *bucket_count = bc;
    *excess_count = ex;
}

// vector<uint32_t> as a stack: the excess-list container (vector_detail.cuh:181-302, 361-368, 454-512),
// values carried as index_t bit patterns
void*
orc_vector_create(std::int64_t capacity)
{
    auto* v = new index_stack();
    v->init(static_cast<index_t>(capacity));
    v->size_ = 0;
    std::memset(v->occ.w, 0, sizeof(unsigned int) * static_cast<std::size_t>((v->cap + 31) / 32));
    return v;
}

void
orc_vector_destroy(void* p)
{
    auto* v = static_cast<index_stack*>(p);
    v->release();
    delete v;
}

std::int64_t orc_vector_size(void* p) { return static_cast<index_stack*>(p)->size(); }

void orc_vector_fill_iota(void* p, std::uint32_t first) { static_cast<index_stack*>(p)->clear_iota(static_cast<index_t>(first)); }

// push n values; ok[i] = pushed
void
orc_vector_push_back(void* p, const std::uint32_t* v, std::int64_t n, std::uint8_t* ok, int threads)
{
    auto* st = static_cast<index_stack*>(p);
#pragma omp parallel for num_threads(threads > 0 ? threads : 1) schedule(static)
    for (std::int64_t i = 0; i < n; ++i)
    {
        const bool r = st->push_back(static_cast<index_t>(v[i]));
        if (ok) ok[i] = r ? 1 : 0;
    }
}

// pop n values from the top; ok[i] = popped, out[i] = value
void
orc_vector_pop_back(void* p, std::uint32_t* out, std::int64_t n, std::uint8_t* ok, int threads)
{
    auto* st = static_cast<index_stack*>(p);
#pragma omp parallel for num_threads(threads > 0 ? threads : 1) schedule(static)
    for (std::int64_t i = 0; i < n; ++i)
    {
        index_t v = 0;
        const bool r = st->pop_back(&v);
        if (ok) ok[i] = r ? 1 : 0;
        out[i] = r ? static_cast<std::uint32_t>(v) : 0xFFFFFFFFU;
    }
}

// elements bottom to top; returns the count
std::int64_t
orc_vector_contents(void* p, std::uint32_t* out)
{
    auto* st = static_cast<index_stack*>(p);
    const index_t n = st->size();
    for (index_t i = 0; i < n; ++i) out[i] = static_cast<std::uint32_t>(st->data[i]);
    return n;
}

void* orc_deque_create(std::int64_t capacity) { return new u32_deque(static_cast<index_t>(capacity)); }
void orc_deque_destroy(void* d) { delete static_cast<u32_deque*>(d); }
std::int64_t orc_deque_size(void* d) { return static_cast<u32_deque*>(d)->size(); }

// push n values; ok[i] = pushed
void
orc_deque_push_back(void* d, const std::uint32_t* v, std::int64_t n, std::uint8_t* ok, int threads)
{
    u32_deque* q = static_cast<u32_deque*>(d);
#pragma omp parallel for num_threads(threads > 0 ? threads : 1) schedule(static)
    for (std::int64_t i = 0; i < n; ++i)
    {
        const bool r = q->push_back(v[i]);
        if (ok) ok[i] = r ? 1 : 0;
    }
}

// pop n values from the front; ok[i] = popped, out[i] = value
void
orc_deque_pop_front(void* d, std::uint32_t* out, std::int64_t n, std::uint8_t* ok, int threads)
{
    u32_deque* q = static_cast<u32_deque*>(d);
#pragma omp parallel for num_threads(threads > 0 ? threads : 1) schedule(static)
    for (std::int64_t i = 0; i < n; ++i)
    {
        unsigned int v = 0;
        const bool r = q->pop_front(&v);
        if (ok) ok[i] = r ? 1 : 0;
        out[i] = r ? v : 0xFFFFFFFFU;
    }
}

// push n values at the front; ok[i] = pushed
void
orc_deque_push_front(void* d, const std::uint32_t* v, std::int64_t n, std::uint8_t* ok, int threads)
{
    u32_deque* q = static_cast<u32_deque*>(d);
#pragma omp parallel for num_threads(threads > 0 ? threads : 1) schedule(static)
    for (std::int64_t i = 0; i < n; ++i)
    {
        const bool r = q->push_front(v[i]);
        if (ok) ok[i] = r ? 1 : 0;
    }
}

// pop n values from the back; ok[i] = popped, out[i] = value
void
orc_deque_pop_back(void* d, std::uint32_t* out, std::int64_t n, std::uint8_t* ok, int threads)
{
    u32_deque* q = static_cast<u32_deque*>(d);
#pragma omp parallel for num_threads(threads > 0 ? threads : 1) schedule(static)
    for (std::int64_t i = 0; i < n; ++i)
    {
        unsigned int v = 0;
        const bool r = q->pop_back(&v);
        if (ok) ok[i] = r ? 1 : 0;
        out[i] = r ? v : 0xFFFFFFFFU;
    }
}

// The voxel caller pattern (a user loop in the reference's applications): per point, pop a slot
// from the free list, insert (key, slot), and push the slot back when the key was already present.
// Sequential, so the result is deterministic.  Returns the number of new keys.
std::int64_t
orc_voxel_insert(void* t, void* d, const std::uint64_t* keys, std::int64_t n)
{
    auto* tab = dynamic_cast<table<std::uint64_t, std::uint32_t>*>(static_cast<table_base*>(t));
    u32_deque* q = static_cast<u32_deque*>(d);
    if (!tab) return -1;
    std::int64_t added = 0;
    for (std::int64_t i = 0; i < n; ++i)
    {
        unsigned int slot = 0;
        if (!q->pop_front(&slot)) continue;
        kv<std::uint64_t, std::uint32_t> v{keys[i], slot};
        if (tab->insert_one(v))
            ++added;
        else
            q->push_back(slot);
    }
    return added;
}

// contents of the deque in front-to-back order; returns the count
std::int64_t
orc_deque_contents(void* d, std::uint32_t* out)
{
    u32_deque* q = static_cast<u32_deque*>(d);
    const index_t n = q->size();
    for (index_t i = 0; i < n; ++i) out[i] = q->data[(q->begin_ + static_cast<unsigned>(i)) % static_cast<unsigned>(q->cap)];
    return n;
}

} // extern "C"
