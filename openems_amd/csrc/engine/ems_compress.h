// Host-side deduplication of per-cell coefficient tuples for the compressed
// device layout (the reference's Operator_SSE_Compressed::CompressOperator,
// FDTD/operator_sse_compressed.cpp:114-175, applied per cell instead of per
// 4-lane z vector).
#pragma once

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace ems {

namespace detail {

// open-addressing set of K-float tuples stored in `keys` (K floats per entry)
template <int K>
struct TupleSet {
    std::vector<float> keys;
    std::vector<uint32_t> slot;   // entry + 1, 0 = empty
    size_t count = 0;

    static uint64_t hash(const float* t)
    {
        uint64_t h = 0x9E3779B97F4A7C15ull;
        for (int k = 0; k < K; ++k) {
            uint32_t b;
            std::memcpy(&b, &t[k], 4);
            if (t[k] == 0.0f) b = 0;   // +0 and -0 are one key
            h ^= b;
            h *= 0xFF51AFD7ED558CCDull;
            h ^= h >> 32;
        }
        return h;
    }
    static bool same(const float* a, const float* b)
    {
        for (int k = 0; k < K; ++k)
            if (!(a[k] == b[k])) return false;
        return true;
    }
    void grow()
    {
        std::vector<uint32_t> old;
        old.swap(slot);
        slot.assign(old.empty() ? 1024 : 2 * old.size(), 0u);
        const size_t mask = slot.size() - 1;
        for (uint32_t e : old)
            if (e) {
                size_t s = hash(&keys[(size_t)(e - 1) * K]) & mask;
                while (slot[s]) s = (s + 1) & mask;
                slot[s] = e;
            }
    }
    // entry number of tuple t, inserting it (as a copy) when new
    uint32_t find_or_insert(const float* t)
    {
        if (2 * (count + 1) > slot.size()) grow();
        const size_t mask = slot.size() - 1;
        size_t s = hash(t) & mask;
        while (slot[s]) {
            const uint32_t e = slot[s] - 1;
            if (same(&keys[(size_t)e * K], t)) return e;
            s = (s + 1) & mask;
        }
        keys.insert(keys.end(), t, t + K);
        slot[s] = (uint32_t)(++count);
        return (uint32_t)(count - 1);
    }
};

template <typename Fn>
void run_chunks(size_t n, unsigned nt, Fn fn)
{
    if (nt <= 1 || n < 4096) {
        fn(0u, (size_t)0, n);
        return;
    }
    std::vector<std::thread> th;
    const size_t per = (n + nt - 1) / nt;
    for (unsigned t = 0; t < nt; ++t) {
        const size_t a = t * per, b = a + per < n ? a + per : n;
        if (a >= b) break;
        th.emplace_back(fn, t, a, b);
    }
    for (auto& t : th) t.join();
}

}  // namespace detail

// Deduplicate `n` tuples of K floats.  Tuple i is get(i, out[K]); tuples are
// visited in index order i = 0..n-1 and numbered in first-seen order.  Two
// tuples are the same when every float compares equal by value (so +0 and -0
// are one key, as with std::map over float operator<); the stored entry is the
// first one seen.  Runs on `nthreads` host threads with a deterministic result.
//   idx[i]  : unique number of tuple i
//   uniq    : K floats per unique tuple, first-seen order
// Returns the number of unique tuples, or 0 if it would exceed max_unique
// (idx / uniq are then unspecified).
template <int K, typename Get>
size_t dedup_tuples(size_t n, Get get, std::vector<uint32_t>& idx, std::vector<float>& uniq, size_t max_unique,
                    unsigned nthreads)
{
    idx.assign(n, 0u);
    uniq.clear();
    if (n == 0) return 0;
    const unsigned nt = nthreads ? nthreads : 1;
    // 1) every chunk numbers its own tuples in first-seen order
    std::vector<detail::TupleSet<K>> local(nt);
    std::vector<char> over(nt, 0);
    detail::run_chunks(n, nt, [&](unsigned t, size_t a, size_t b) {
        float tup[K];
        auto& S = local[t];
        for (size_t i = a; i < b; ++i) {
            get(i, tup);
            idx[i] = S.find_or_insert(tup);
            if (S.count > max_unique) {
                over[t] = 1;
                return;
            }
        }
    });
    for (char o : over)
        if (o) return 0;
    // 2) merge chunk tables in chunk order: global first-seen order
    detail::TupleSet<K> G;
    std::vector<std::vector<uint32_t>> remap(nt);
    for (unsigned t = 0; t < nt; ++t) {
        const auto& S = local[t];
        remap[t].resize(S.count);
        for (size_t e = 0; e < S.count; ++e) {
            remap[t][e] = G.find_or_insert(&S.keys[e * K]);
            if (G.count > max_unique) return 0;
        }
    }
    // 3) local -> global numbers
    detail::run_chunks(n, nt, [&](unsigned t, size_t a, size_t b) {
        const auto& R = remap[t];
        for (size_t i = a; i < b; ++i) idx[i] = R[idx[i]];
    });
    uniq.swap(G.keys);
    return G.count;
}


// Per-cell compression of the N-I-J-K operator (vv, vi, ii, iv, [3][Nx][Ny][Nz]).
// Cells are numbered x -> y -> z (x outermost, the order of the reference's
// compression loop).  Outputs, for the device:
//   idx_soa : uint16 per cell in the x-fastest layout, length cs (padding 0)
//   tabV    : 8 floats per unique tuple {vv_x, vi_x, vv_y, vi_y, vv_z, vi_z, 0, 0}
//   tabI    : 8 floats per unique tuple {ii_x, iv_x, ii_y, iv_y, ii_z, iv_z, 0, 0}
// Returns the number of unique tuples, or 0 if more than 65536.
size_t compress_cells(const unsigned N[3], const float* vv, const float* vi, const float* ii, const float* iv,
                      unsigned long long cs, std::vector<uint16_t>& idx_soa, std::vector<float>& tabV,
                      std::vector<float>& tabI, unsigned nthreads);

// Rebuild the N-I-J-K operator arrays from the reference's SSE-compressed form
// (Operator_SSE_Compressed: m_Op_index [Nx][Ny][Nz] with only z < nV used,
// f4_*_Compressed[3][U] f4vectors whose lane l holds z = l*nV + v).
void expand_sse_compressed(const unsigned N[3], const unsigned* op_index, unsigned U, const float* const vv4[3],
                           const float* const vi4[3], const float* const ii4[3], const float* const iv4[3],
                           float* vv, float* vi, float* ii, float* iv, unsigned nthreads);

unsigned host_threads();

}  // namespace ems
