// gfx950 kernels and device runtime of the de Pina support-vector phase loop.
//
// Data layout in HBM (native form)
//   trees      parent_pos u32[P], pos_coord i32[P]: every root's shortest-path tree in BFS-level
//              order (parent before child), one contiguous segment per root (tree_off i64[F+1]).
//              pos_coord holds the cycle-space coordinate of the edge to the parent (-1: none),
//              so the parity pass needs no gather through the 2m'-entry coordinate map.
//              pos_pre/pos_end u32[P] hold DFS pre-order intervals (ancestor tests for C_e).
//   candidates rank-indexed (cand_*) for the cycle vector, and root-major copies ordered by rank
//              (rc_*) for the search, 16 B per candidate.
//   support    word-major matrix Sw u64[W][Npad]: Sw[w][j] = word w of S_j without its own bit j
//              (bit j of S_j is always 1).  A phase-e update touches words 0..e/64 of columns
//              j > e, coalesced across j.  S_e itself is also kept contiguous for the parity pass.
//
// One kernel per phase e (k_phase)
//   prologue   the update (c), one thread per column j > e: the dot product of phase e-1
//              (p_j = <C_{e-1}, S_j>, from C_{e-1}'s nonzero words below e-1) and the XOR of phase
//              e-2 (S_j ^= S_{e-2} where phase e-2's p_j was 1, over S_{e-2}'s nonzero words).
//              Both steps of a column run in one thread, so no column is read while it is written.
//   body       (a)+(b): tree-path parity of every root with S_e in LDS, kept as one bit per position
//              in LDS, then the root's candidates in rank order until the first odd one ->
//              atomicMin into best[e].  Trees larger than 2048 positions are shared by the 8 waves of
//              a block; smaller ones take one wave each.
//   epilogue   the last block to finish builds C_e for the chosen candidate (ancestor test against
//              DFS pre-order intervals over the root's tree), C_e's nonzero words below e, and
//              column e+1 of (c), which makes S_{e+1} final and writes it contiguously.
// S_e rotates through four contiguous buffers and C_e through two, so a phase reads earlier
// phases' outputs while writing its own.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <functional>
#include <mutex>
#include <numeric>
#include <queue>
#include <string>
#include <thread>
#include <vector>

#include "mcb_hip.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define HIP_TRY(call)                                                                            \
    do {                                                                                         \
        hipError_t _e = (call);                                                                  \
        if (_e != hipSuccess)                                                                    \
            return fail(MCB_E_DEVICE, std::string(#call) + ": " + hipGetErrorString(_e));       \
    } while (0)

// The reference names abort on a runtime error like CudaError (include/utils.h:36-47).
#define HIP_OR_EXIT(call)                                                                        \
    do {                                                                                         \
        hipError_t _e = (call);                                                                  \
        if (_e != hipSuccess) {                                                                  \
            std::fprintf(stderr, "Error :%s : %s : line No = %d\n", hipGetErrorString(_e),       \
                         __FILE__, __LINE__);                                                    \
            std::exit(static_cast<int>(_e));                                                     \
        }                                                                                        \
    } while (0)

constexpr uint32_t kNoParent = 0xFFFFFFFFu;
constexpr uint32_t kNoRank = 0xFFFFFFFFu;
constexpr int kWave = 64;
constexpr int kLdsLimit = 64 * 1024;
#ifndef MCB_UNROLL
#define MCB_UNROLL 8
#endif
#ifndef MCB_BLOCK_WAVES
#define MCB_BLOCK_WAVES 8
#endif
#ifndef MCB_BIG_WINDOWS
#define MCB_BIG_WINDOWS 32
#endif
constexpr int kUnroll = MCB_UNROLL;           // tree windows per pipelined load group
#ifndef MCB_TW_UNROLL
#define MCB_TW_UNROLL 2
#endif
constexpr int kTwUnroll = MCB_TW_UNROLL;      // windows per group of a one-wave tree walk
constexpr int kBlockWaves = MCB_BLOCK_WAVES;  // waves per block of the parity/search kernel
constexpr uint32_t kBigPositions = 64u * MCB_BIG_WINDOWS;  // larger trees are shared by a whole block

// Index of the calling wave within its workgroup, in a scalar register: the compiler cannot prove
// threadIdx.x / 64 wave-uniform, and loop bounds derived from it would make everything they carry
// divergent.
__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / kWave); }

// A value every lane holds, moved to scalar registers (a vector load from a uniform address is
// otherwise taken as divergent, and so is everything computed from it).
__device__ __forceinline__ int64_t uniform64(int64_t x) {
    const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(static_cast<uint32_t>(x))));
    const uint32_t hi =
        static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(static_cast<uint32_t>(static_cast<uint64_t>(x) >> 32))));
    return static_cast<int64_t>((static_cast<uint64_t>(hi) << 32) | lo);
}
__device__ __forceinline__ uint32_t uniform32(uint32_t x) {
    return static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(x)));
}

// The loads below are unconditional (a clamped address, the value selected afterwards): a load
// under a divergent branch makes the compiler wait for every outstanding load at the join, which
// would serialise the prefetch pipeline of the tree kernels.
__device__ __forceinline__ uint32_t bit_of(const uint64_t *s, int32_t c) {
    const uint32_t cc = c >= 0 ? static_cast<uint32_t>(c) : 0u;
    const uint32_t bit = static_cast<uint32_t>((s[cc >> 6] >> (cc & 63)) & 1ull);
    return c >= 0 ? bit : 0u;
}

// Parents and coordinates of kUnroll windows, loaded raw (clamped address, no select): the
// values stay in flight until the group is processed.  Buffers alternate between two groups so
// that no register holding an in-flight load is copied (a copy would wait for the load).
template <int U>
struct WindowGroupT {
    uint32_t par[U];
    int32_t co[U];
};
using WindowGroup = WindowGroupT<kUnroll>;

template <int U>
__device__ __forceinline__ void load_group(WindowGroupT<U> &g, const uint32_t *pp, const int32_t *pc, uint32_t npos,
                                           uint32_t first_pos, uint32_t stride) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const uint32_t pos = first_pos + u * stride;
        const uint32_t at = pos < npos ? pos : 0u;
        g.par[u] = pp[at];
        g.co[u] = pc[at];
    }
}

__device__ __forceinline__ uint32_t pbit(const uint64_t *p, uint32_t pos) {
    return static_cast<uint32_t>((p[pos >> 6] >> (pos & 63)) & 1ull);
}

// Lanes of one window whose parent lies in the same window (the window straddles a level
// boundary): settle them by repeated ballots; each round settles at least the lowest pending lane.
__device__ __forceinline__ uint32_t settle_in_window(uint32_t val, bool pending_lane, uint32_t p, uint32_t wb,
                                                     uint32_t sb) {
    bool resolved = !pending_lane;
    uint64_t pending = __ballot(!resolved);
    while (pending) {
        const uint64_t rm = __ballot(resolved);
        const uint64_t pm = __ballot(val);
        if (!resolved) {
            const uint32_t k = p - wb;
            if ((rm >> k) & 1ull) {
                val = sb ^ static_cast<uint32_t>((pm >> k) & 1ull);
                resolved = true;
            }
        }
        pending = __ballot(!resolved);
    }
    return val;
}

// Parity of every position of one tree segment into the wave's LDS bit array P (bit p = parity
// of position p).  Positions are taken 64 at a time (a window, one lane each); the parents and
// coordinates of the next kUnroll windows are loaded while the current kUnroll windows are
// resolved.  A parent before the window is read from P (already final); a parent inside the
// window only occurs where the window straddles a level boundary and is settled by ballots.
// One window of tree_parity: par/co are the raw loads for position wb + lane.
template <typename WriteFn>
__device__ __forceinline__ void small_window(uint32_t wb, uint32_t npos, uint32_t rawp, int32_t rawc,
                                             const uint64_t *S, uint64_t *P, int lane, WriteFn &write) {
    const bool v = wb + lane < npos;
    const uint32_t p = v ? rawp : kNoParent;
    const uint32_t sb = bit_of(S, v ? rawc : -1);
    const bool before = p < wb;
    const uint32_t pb = pbit(P, before ? p : 0u);
    uint32_t val = sb ^ (before ? pb : 0u);
    const bool inwin = (p != kNoParent) && (p >= wb);
    if (__ballot(inwin)) val = settle_in_window(val, inwin, p, wb, sb);
    const uint64_t word = __ballot(val);
    if (lane == 0) P[wb / kWave] = word;
    write(wb + lane, val);
}

template <typename WriteFn>
__device__ __forceinline__ void small_group(const WindowGroup &g, uint32_t base, uint32_t npos, const uint64_t *S,
                                            uint64_t *P, int lane, WriteFn &write) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
        const uint32_t wb = base + u * kWave;
        if (wb >= npos) break;
        small_window(wb, npos, g.par[u], g.co[u], S, P, lane, write);
    }
}

// Parity of every position of one tree segment into the wave's LDS bit array P (bit p = parity
// of position p).  Positions are taken 64 at a time (a window, one lane each); the parents and
// coordinates of the next kUnroll windows are loaded while the current kUnroll windows are
// resolved.  A parent before the window is read from P (already final); a parent inside the
// window only occurs where the window straddles a level boundary and is settled by ballots.
template <typename WriteFn>
__device__ __forceinline__ void tree_parity(const uint32_t *pp, const int32_t *pc, uint32_t npos, const uint64_t *S,
                                            uint64_t *P, int lane, WriteFn write) {
    constexpr uint32_t G = kWave * kUnroll;
    WindowGroup A, B;
    load_group(A, pp, pc, npos, lane, kWave);
    for (uint32_t base = 0; base < npos; base += 2 * G) {
        load_group(B, pp, pc, npos, base + G + lane, kWave);
        small_group(A, base, npos, S, P, lane, write);
        if (base + G >= npos) break;
        load_group(A, pp, pc, npos, base + 2 * G + lane, kWave);
        small_group(B, base + G, npos, S, P, lane, write);
    }
}

// Parity of one big root's tree by all waves of a workgroup.  Window w (positions 64w..64w+63)
// belongs to wave w mod nw; each wave walks its windows in increasing order with the next group of
// loads in flight.  A parent in an earlier window is read from LDS once that window's ready stamp
// equals this root's generation; every dependency points to a lower window and the lowest unfinished
// window never waits on another wave, so the waits always end.
// Parity of one window in a reference-layout tree (used by the chunked drop-in kernel): parents
// before wb are final in P; parents inside the window are settled with ballots.
__device__ __forceinline__ uint32_t window_parity(uint64_t *P, uint32_t wb, uint32_t par, uint32_t b, int lane) {
    const bool before = par < wb;
    const uint32_t pb = pbit(P, before ? par : 0u);
    uint32_t val = b ^ (before ? pb : 0u);
    const bool inwin = (par != kNoParent) && (par >= wb);
    if (__ballot(inwin)) val = settle_in_window(val, inwin, par, wb, b);
    const uint64_t word = __ballot(val);
    if (lane == 0) P[wb >> 6] = word;
    __builtin_amdgcn_wave_barrier();
    return val;
}

template <typename WriteFn>
__device__ __forceinline__ void coop_group(const WindowGroup &g, uint32_t w0, uint32_t nwin, uint32_t npos,
                                           const uint64_t *S, uint64_t *P, uint32_t *ready, uint32_t gen, int nw,
                                           int lane, WriteFn &write) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
        const uint32_t w = w0 + u * nw;
        if (w >= nwin) break;
        const uint32_t wb = w * kWave;
        const bool v = wb + lane < npos;
        const uint32_t p = v ? g.par[u] : kNoParent;
        const uint32_t sb = bit_of(S, v ? g.co[u] : -1);
        const bool before = p < wb;
        const uint32_t pw = before ? p / kWave : 0u;
        // wait until the parent's window is final (windows of this wave are final already)
        while (before && __hip_atomic_load(ready + pw, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) != gen)
            __builtin_amdgcn_s_sleep(1);
        const uint32_t pb = pbit(P, before ? p : 0u);
        uint32_t val = sb ^ (before ? pb : 0u);
        const bool inwin = (p != kNoParent) && (p >= wb);
        if (__ballot(inwin)) val = settle_in_window(val, inwin, p, wb, sb);
        const uint64_t word = __ballot(val);
        if (lane == 0) {
            P[w] = word;
            __hip_atomic_store(ready + w, gen, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
        write(wb + lane, val);
    }
}

// Parity of one big root's tree by all waves of a workgroup.  Window w (positions 64w..64w+63)
// belongs to wave w mod nw; each wave walks its windows in increasing order with the next group of
// loads in flight.  A parent in an earlier window is read from LDS once that window's ready stamp
// equals this root's generation; every dependency points to a lower window and the lowest unfinished
// window never waits on another wave, so the waits always end.
template <typename WriteFn>
__device__ __forceinline__ void tree_parity_coop(const uint32_t *pp, const int32_t *pc, uint32_t npos,
                                                 const uint64_t *S, uint64_t *P, uint32_t *ready, uint32_t gen,
                                                 int wid, int nw, int lane, WriteFn write) {
    const uint32_t nwin = (npos + kWave - 1) / kWave;
    const uint32_t stride = static_cast<uint32_t>(nw) * kWave;  // positions between a wave's windows
    const uint32_t gstep = kUnroll * static_cast<uint32_t>(nw);   // windows between a wave's groups
    WindowGroup A, B;
    load_group(A, pp, pc, npos, wid * kWave + lane, stride);
    for (uint32_t w0 = wid; w0 < nwin; w0 += 2 * gstep) {
        load_group(B, pp, pc, npos, (w0 + gstep) * kWave + lane, stride);
        coop_group(A, w0, nwin, npos, S, P, ready, gen, nw, lane, write);
        if (w0 + gstep >= nwin) break;
        load_group(A, pp, pc, npos, (w0 + 2 * gstep) * kWave + lane, stride);
        coop_group(B, w0 + gstep, nwin, npos, S, P, ready, gen, nw, lane, write);
    }
}

struct PhaseArgs {
    // trees, in BFS-level order per root, with DFS pre-order intervals for ancestor tests
    const int64_t *tree_off;
    const uint32_t *parent_pos;
    const int32_t *pos_coord;
    const uint32_t *pos_pre, *pos_end;  // x is an ancestor-or-self of y iff pre[x] <= pre[y] < end[x]
    const uint32_t *big_list;           // roots with more than kBigPositions positions, largest first
    const uint32_t *small_list;         // the other roots, largest first
    const uint32_t *root_min_rank;      // lowest candidate rank of each root
    // candidates: root-major (search) and rank-indexed (cycle vector)
    const int64_t *rc_off;
    const uint32_t *rc_pu, *rc_pv, *rc_rank;
    const int32_t *rc_coord;
    const uint32_t *cand_root, *cand_pu, *cand_pv;
    const int32_t *cand_coord;
    // support vectors: word-major matrix without the diagonal, plus contiguous copies in a ring of
    // four (S_{e-2}, S_{e-1}, S_e, S_{e+1}) with the lists of their nonzero words
    uint64_t *Sw;
    int64_t npad;
    const uint64_t *S;                     // S_e
    const uint64_t *S_m1;                  // S_{e-1}
    const uint64_t *S_m2;                  // S_{e-2}
    uint64_t *S_next;                      // S_{e+1}, written by the epilogue
    const uint32_t *E_m2, *Ecount_m2;      // nonzero words of S_{e-2}
    uint32_t *E_next, *Ecount_next;        // nonzero words of S_{e+1}, written by the epilogue
    // cycle vectors: C_{e-1} and its nonzero words below e-1 (read), C_e (written)
    const uint64_t *C_prev;
    const uint32_t *D_prev, *Dcount_prev;
    const uint32_t *c_prev;                // <C_{e-1}, S_{e-2}>
    uint64_t *C;
    uint32_t *D, *Dcount;
    uint32_t *c_out;                       // <C_e, S_{e-1}>, written by the epilogue
    uint64_t *C_store;                     // optional: row e of the stored cycle vectors
    uint8_t *flag;           // per column j: p_j of the latest phase whose dot products are done
    uint32_t *best;          // &best[e]
    uint32_t *done;          // &done[e]: blocks finished
    uint32_t *status;
    uint8_t *parity_out;     // parity-only mode: one byte per position
    uint64_t *wave_stats;    // optional: per wave, positions streamed, candidates read, update bytes
    int32_t e, N, W, n_big, n_small, p_words_big, p_words_small;
    int32_t upd_tiles;  // prologue tiles: 64 columns from e+1 each (0 in phase 0)
};


// Scans candidate windows k0, k0 + stride, ... of one root (ranks ascending within the root) and
// lowers best to the rank of the first odd one; stops early once the ranks cannot win.
__device__ __forceinline__ void scan_candidates(const PhaseArgs &a, const uint64_t *Sx, const uint64_t *P,
                                                int64_t first, int64_t ce, int64_t stride, uint32_t bestk, int lane,
                                                uint64_t &n_cand) {
    for (int64_t k0 = first; k0 < ce; k0 += stride) {
        const int64_t k = k0 + lane;
        const bool v = k < ce;
        n_cand += static_cast<uint64_t>(ce - k0 < kWave ? ce - k0 : kWave);
        const uint32_t rk = v ? a.rc_rank[k] : kNoRank;
        const bool live = rk < bestk;
        uint32_t odd = 0;
        if (live) odd = bit_of(Sx, a.rc_coord[k]) ^ pbit(P, a.rc_pu[k]) ^ pbit(P, a.rc_pv[k]);
        const uint64_t m = __ballot(odd != 0);
        if (m) {
            const int firstlane = __ffsll(static_cast<unsigned long long>(m)) - 1;
            const uint32_t win = __shfl(rk, firstlane);
            if (lane == 0) atomicMin(a.best, win);
            return;
        }
        if (__ballot(v && !live)) return;
        bestk = __hip_atomic_load(a.best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// The support-vector update (c) of phase p, S_j ^= S_p for every j > p with <C_p, S_j> = 1, is
// split across two later kernels so that each column is only ever touched by one thread:
//   phase p+1's prologue computes f_j = <C_p, S_j> for j > p+1 (flag[j]);
//   phase p+2's prologue applies S_j ^= S_p where f_j = 1 (j > p+2);
//   column p+1 (resp. p+2) is finished by phase p's (resp. p+1's) epilogue instead.
// In phase e's prologue the thread of column j > e therefore first reads f_j of phase e-2, computes
// <C_{e-1}, S_j> on the column as it still is, corrects it for the pending XOR of S_{e-2}
// (<C_{e-1}, S_j ^ S_{e-2}> = <C_{e-1}, S_j> ^ <C_{e-1}, S_{e-2}>), stores it as f_j of phase e-1,
// and then applies the pending XOR.  The dot product needs only C_{e-1}'s nonzero words below e-1
// (S_j has no other bits there besides its own bit j, which meets C_{e-1}[j]); the XOR needs only
// S_{e-2}'s nonzero words.  Loads are issued in batches before the stores.
// Returns whether the column took the pending XOR.
__device__ __forceinline__ uint32_t column_update(const PhaseArgs &a, int64_t j) {
    constexpr int kB = 8;
    const uint32_t f_old = a.e >= 2 ? a.flag[j] : 0u;
    uint32_t p = static_cast<uint32_t>((a.C_prev[j >> 6] >> (j & 63)) & 1ull) ^ (f_old & *a.c_prev);
    const uint32_t nd = *a.Dcount_prev;
#pragma unroll 4
    for (uint32_t q = 0; q < nd; ++q) {
        const uint32_t w = a.D_prev[q];
        p ^= static_cast<uint32_t>(__popcll(a.C_prev[w] & a.Sw[static_cast<int64_t>(w) * a.npad + j]) & 1);
    }
    a.flag[j] = static_cast<uint8_t>(p);
    if (!f_old) return 0u;
    const uint32_t ne = *a.Ecount_m2;
    for (uint32_t q0 = 0; q0 < ne; q0 += kB) {
        uint32_t w[kB];
        uint64_t se[kB], v[kB];
#pragma unroll
        for (int k = 0; k < kB; ++k) {
            const uint32_t q = q0 + k;
            w[k] = a.E_m2[q < ne ? q : 0];
            se[k] = q < ne ? a.S_m2[w[k]] : 0ull;
        }
#pragma unroll
        for (int k = 0; k < kB; ++k) v[k] = a.Sw[static_cast<int64_t>(w[k]) * a.npad + j];
#pragma unroll
        for (int k = 0; k < kB; ++k)
            if (se[k]) a.Sw[static_cast<int64_t>(w[k]) * a.npad + j] = v[k] ^ se[k];
    }
    return 1u;
}

// Epilogue of phase e, run by the last block to finish:
//   C_e of the chosen candidate: coordinates of the edges on root -> u, root -> v and of the
//   candidate's own edge (the two paths meet only at the root, so setting bits equals XOR), found by
//   testing every position of the root's tree against the DFS intervals of u and v;
//   the list of C_e's nonzero words below e, and <C_e, S_{e-1}> for the next prologue;
//   column e+1: the pending XOR of S_{e-1}, then phase e's own update, which makes S_{e+1} final;
//   S_{e+1} and the list of its nonzero words, contiguous, for phase e+1.
// LDS: C_e (W words) then the word list (W uint32).
__device__ void phase_epilogue(const PhaseArgs &a, uint64_t *lds) {
    constexpr int kB = 8;
    __shared__ uint32_t cnt;
    __shared__ uint32_t par;
    __shared__ uint32_t cpar;
    const int tid = threadIdx.x;
    uint64_t *Cl = lds;
    uint32_t *Dl = reinterpret_cast<uint32_t *>(lds + a.W);
    const uint32_t rank = __hip_atomic_load(a.best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (int i = tid; i < a.W; i += blockDim.x) Cl[i] = 0;
    if (tid == 0) {
        cnt = 0;
        par = 0;
        cpar = 0;
    }
    __syncthreads();
    if (rank != kNoRank) {
        const uint32_t r = a.cand_root[rank];
        const int64_t off = a.tree_off[r];
        const uint32_t npos = static_cast<uint32_t>(a.tree_off[r + 1] - off);
        const uint32_t pu = a.pos_pre[off + a.cand_pu[rank]];
        const uint32_t pv = a.pos_pre[off + a.cand_pv[rank]];
        if (tid == 0) {
            const int32_t c = a.cand_coord[rank];
            if (c >= 0) atomicOr(reinterpret_cast<unsigned long long *>(&Cl[c >> 6]), 1ull << (c & 63));
        }
        for (uint32_t x0 = 1 + tid; x0 < npos; x0 += kB * blockDim.x) {
            uint32_t lo[kB], hi[kB];
            int32_t c[kB];
#pragma unroll
            for (int k = 0; k < kB; ++k) {
                const uint32_t x = x0 + k * blockDim.x;
                const uint32_t at = x < npos ? x : 0u;
                lo[k] = a.pos_pre[off + at];
                hi[k] = a.pos_end[off + at];
                c[k] = x < npos ? a.pos_coord[off + at] : -1;
            }
#pragma unroll
            for (int k = 0; k < kB; ++k) {
                const bool on = (lo[k] <= pu && pu < hi[k]) || (lo[k] <= pv && pv < hi[k]);
                if (on && c[k] >= 0)
                    atomicOr(reinterpret_cast<unsigned long long *>(&Cl[c[k] >> 6]), 1ull << (c[k] & 63));
            }
        }
    } else if (tid == 0) {
        atomicOr(a.status, 1u);
    }
    __syncthreads();
    const int32_t dlim = a.e > 0 ? ((a.e - 1) >> 6) + 1 : 0;
    for (int i = tid; i < a.W; i += blockDim.x) {
        const uint64_t v = Cl[i];
        a.C[i] = v;
        if (a.C_store) a.C_store[i] = v;
        if (i < dlim && v != 0) {
            const uint32_t slot = atomicAdd(&cnt, 1u);
            Dl[slot] = static_cast<uint32_t>(i);
            a.D[slot] = static_cast<uint32_t>(i);
        }
    }
    __syncthreads();
    const uint32_t nd = cnt;
    const int64_t j = static_cast<int64_t>(a.e) + 1;
    const bool have_col = j < a.N;
    // column e+1 after the pending XOR of S_{e-1} (its flag was set by this phase's prologue)
    const uint32_t f_old = (a.e >= 1 && have_col) ? a.flag[j] : 0u;
    uint32_t part = 0, cp = 0;
    for (uint32_t q = tid; q < nd; q += blockDim.x) {
        const uint32_t w = Dl[q];
        if (have_col) {
            uint64_t col = a.Sw[static_cast<int64_t>(w) * a.npad + j];
            if (f_old) col ^= a.S_m1[w];
            part ^= static_cast<uint32_t>(__popcll(Cl[w] & col) & 1);
        }
        if (a.e >= 1) cp ^= static_cast<uint32_t>(__popcll(Cl[w] & a.S_m1[w]) & 1);
    }
    if (part) atomicXor(&par, 1u);
    if (cp) atomicXor(&cpar, 1u);
    __syncthreads();
    if (tid == 0) {
        *a.Dcount = nd;
        *a.c_out = cpar;
        cnt = 0;
    }
    if (!have_col) return;
    const uint32_t p = par ^ static_cast<uint32_t>((Cl[j >> 6] >> (j & 63)) & 1ull);
    const int32_t wl = static_cast<int32_t>(j >> 6) + 1;
    const int32_t we = a.e >> 6;
    const int32_t wm1 = a.e >= 1 ? ((a.e - 1) >> 6) : -1;
    __syncthreads();
    for (int32_t w = tid; w < wl; w += blockDim.x) {
        uint64_t *cell = a.Sw + static_cast<int64_t>(w) * a.npad + j;
        const uint64_t old = *cell;
        uint64_t v = old;
        if (f_old && w <= wm1) v ^= a.S_m1[w];
        if (p && w <= we) v ^= a.S[w];
        if (v != old) *cell = v;
        if (w == static_cast<int32_t>(j >> 6)) v |= 1ull << (j & 63);
        a.S_next[w] = v;
        if (v) a.E_next[atomicAdd(&cnt, 1u)] = static_cast<uint32_t>(w);
    }
    __syncthreads();
    if (tid == 0) *a.Ecount_next = cnt;
}

// One phase e: the column updates of (c) (dot products of phase e-1, XOR of phase e-2) on columns
// > e, then (a)+(b), then the epilogue in the last block.
// (a)+(b): block b takes big roots b, b + grid, ... (all waves of the block on one tree), then wave
// w of the grid takes small roots w, w + waves, ...; both lists are sorted by size, largest first, so
// the static round-robin spreads the large trees evenly (a shared work counter costs more than it
// balances).  A root whose lowest candidate rank is not below the best rank found so far is skipped.
// SEARCH = false computes the parity of every position of every root into parity_out instead.
template <bool S_IN_LDS, bool SEARCH>
__global__ __launch_bounds__(kBlockWaves * kWave) void k_phase(PhaseArgs a) {
    extern __shared__ uint64_t lds[];
    __shared__ uint32_t skip;
    __shared__ uint32_t last;
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = wave_id();
    const int nw = blockDim.x / kWave;
    uint64_t n_upd = 0;  // algorithmic bytes of the update: 8 per dot word, 16 per XORed word, 2 per flag
    if (SEARCH && a.upd_tiles > 0) {
        // 64-column tiles of the update, handed out from the last wave of the grid backwards
        const int waves = static_cast<int>(gridDim.x) * nw;
        const int gw = static_cast<int>(blockIdx.x) * nw + wid;
        for (int t = waves - 1 - gw; t < a.upd_tiles; t += waves) {
            const int64_t j = static_cast<int64_t>(a.e) + 1 + static_cast<int64_t>(t) * kWave + lane;
            const uint32_t x = j < a.N ? column_update(a, j) : 0u;
            if (a.wave_stats) {
                const uint64_t cols = __popcll(__ballot(j < a.N));
                const uint64_t xors = __popcll(__ballot(x != 0));
                n_upd += cols * (2ull + 8ull * *a.Dcount_prev) + (xors ? xors * 16ull * *a.Ecount_m2 : 0ull);
            }
        }
    }
    const uint64_t *Sx = a.S;
    uint64_t *Pb = lds + (S_IN_LDS ? a.W : 0);
    uint32_t *ready = reinterpret_cast<uint32_t *>(Pb + a.p_words_big);
    uint64_t *Ps = Pb + a.p_words_big + (a.p_words_big + 1) / 2 + static_cast<size_t>(wid) * a.p_words_small;
    if (S_IN_LDS)
        for (int i = threadIdx.x; i < a.W; i += blockDim.x) lds[i] = a.S[i];
    for (int i = threadIdx.x; i < a.p_words_big; i += blockDim.x) ready[i] = 0;
    __syncthreads();
    if (S_IN_LDS) Sx = lds;
    uint64_t n_pos = 0, n_cand = 0;
    uint32_t gen = 0;
    for (uint32_t it = blockIdx.x; it < static_cast<uint32_t>(a.n_big); it += gridDim.x) {
        if (threadIdx.x == 0)
            skip = SEARCH ? (a.root_min_rank[a.big_list[it]] >=
                             __hip_atomic_load(a.best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                          : 0u;
        __syncthreads();
        const uint32_t sk = skip;
        __syncthreads();
        if (sk) continue;
        const uint32_t r = a.big_list[it];
        const int64_t off = a.tree_off[r];
        const uint32_t npos = static_cast<uint32_t>(a.tree_off[r + 1] - off);
        ++gen;
        if (lane == 0 && wid == 0) n_pos += npos;
        if (SEARCH) {
            tree_parity_coop(a.parent_pos + off, a.pos_coord + off, npos, Sx, Pb, ready, gen, wid, nw, lane,
                             [](uint32_t, uint32_t) {});
        } else {
            uint8_t *out = a.parity_out + off;
            tree_parity_coop(a.parent_pos + off, a.pos_coord + off, npos, Sx, Pb, ready, gen, wid, nw, lane,
                             [&](uint32_t pos, uint32_t v) {
                                 if (pos < npos) out[pos] = static_cast<uint8_t>(v);
                             });
        }
        __syncthreads();
        if (SEARCH) {
            const uint32_t bestk = __hip_atomic_load(a.best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            scan_candidates(a, Sx, Pb, a.rc_off[r] + static_cast<int64_t>(wid) * kWave, a.rc_off[r + 1],
                            static_cast<int64_t>(nw) * kWave, bestk, lane, n_cand);
        }
        __syncthreads();
    }
    for (uint32_t it = blockIdx.x * nw + wid; it < static_cast<uint32_t>(a.n_small); it += gridDim.x * nw) {
        const uint32_t r = a.small_list[it];
        uint32_t bestk = kNoRank;
        if (SEARCH) {
            bestk = __hip_atomic_load(a.best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (a.root_min_rank[r] >= bestk) continue;
        }
        const int64_t off = a.tree_off[r];
        const uint32_t npos = static_cast<uint32_t>(a.tree_off[r + 1] - off);
        if (lane == 0) n_pos += npos;
        if (SEARCH) {
            tree_parity(a.parent_pos + off, a.pos_coord + off, npos, Sx, Ps, lane, [](uint32_t, uint32_t) {});
            scan_candidates(a, Sx, Ps, a.rc_off[r], a.rc_off[r + 1], kWave, bestk, lane, n_cand);
        } else {
            uint8_t *out = a.parity_out + off;
            tree_parity(a.parent_pos + off, a.pos_coord + off, npos, Sx, Ps, lane, [&](uint32_t pos, uint32_t v) {
                if (pos < npos) out[pos] = static_cast<uint8_t>(v);
            });
        }
    }
    if (a.wave_stats && lane == 0) {  // each wave owns its slot; launches are stream-ordered
        const size_t gw = static_cast<size_t>(blockIdx.x) * nw + wid;
        a.wave_stats[3 * gw] += n_pos;
        a.wave_stats[3 * gw + 1] += n_cand;
        a.wave_stats[3 * gw + 2] += n_upd;
    }
    if (!SEARCH) return;
    // The last block to finish sees every other block's result for this phase.
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        last = atomicAdd(a.done, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    phase_epilogue(a, lds);
}

// Reference-compatible parity of one chunk: every position of every valid root slot, 0/1 as
// uint32 (gpu/kernel_init_edges.cu:32-75 followed by gpu/kernel_multi_search.cu:62-108).
__global__ __launch_bounds__(256) void k_compat_parity(const int32_t *__restrict__ nt, const int32_t *__restrict__ eo,
                                                       const uint32_t *__restrict__ cols, uint32_t *__restrict__ out,
                                                       const uint64_t *__restrict__ S, int32_t n, int32_t roots,
                                                       int32_t p_words) {
    extern __shared__ uint64_t lds[];
    const int lane = threadIdx.x & (kWave - 1);
    const int wib = wave_id();
    const int wpb = blockDim.x / kWave;
    uint64_t *P = lds + static_cast<size_t>(wib) * p_words;
    const int slot = blockIdx.x * wpb + wib;
    if (slot >= roots) return;
    const int32_t *e = eo + static_cast<int64_t>(slot) * n;
    const uint32_t *c = cols + static_cast<int64_t>(slot) * n;
    uint32_t *o = out + static_cast<int64_t>(slot) * n;
    for (uint32_t base = 0; base < static_cast<uint32_t>(n); base += kWave) {
        const uint32_t pos = base + lane;
        const bool v = pos < static_cast<uint32_t>(n);
        const int32_t edge = v ? e[pos] : -1;
        const uint32_t par = v ? c[pos] : kNoParent;
        const int32_t coord = edge >= 0 ? nt[edge] : -1;
        const uint32_t val = window_parity(P, base, par, bit_of(S, coord), lane);
        if (v) o[pos] = val;
    }
}

__global__ void k_fill_u32(uint32_t *p, uint32_t v, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_copy_u64(const uint64_t *__restrict__ src, uint64_t *__restrict__ dst, int64_t n) {
    int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}

// ---------------------------------------------------------------------------------------------
// Blocked phase loop: up to 64 phases resolved per pass over the trees
//
// Parity is linear in S.  At the start of a block of kk phases e..e+kk-1 the vectors
// T_l = S_{e+l} (all updates of earlier phases applied) are packed into one word per coordinate,
// Tt[c] bit l = T_l[c].  One pass over every tree gives each candidate k the word
// Qr[k] bit l = <C_k, T_l>.  The final vector of phase e+i is S_{e+i} = sum over l of M[i][l] T_l
// (M lower unitriangular, row i = mask m_i), so candidate k is odd in phase e+i exactly when
// popcount(m_i & Qr[k]) is odd.  The block's winners are then the first kk candidates, in rank
// order, whose words are linearly independent over GF(2), and the phase a winner serves is the
// lowest set bit of its word after elimination by the winners before it.  That is a greedy basis,
// computed per rank chunk in parallel and merged chunk group by chunk group (the greedy basis of a
// concatenation is the greedy basis of the concatenated greedy bases).  The later vectors
// j >= e+kk then take the whole block's update at once: b_ji = <C_{e+i}, S_j> xor
// sum_{l<i} b_jl <C_{e+i}, S_{e+l}>, S_j ^= sum_i b_ji S_{e+i}.
// ---------------------------------------------------------------------------------------------

constexpr int kBlockK = 64;      // phases per block (word width)
constexpr int kQWaves = 16;      // waves per workgroup of the tree pass
constexpr int kChunk = 4096;     // ranks per first-level greedy basis (one candidate-bitmap word per lane)
constexpr int kMaxRounds = 16;   // bitmap rounds of 64 words per first-level basis of the blocked form
constexpr int64_t kXorWorkgroups = 2048;  // grid of k_block_xor (eight workgroups per CU)
constexpr int64_t kDotWorkgroups = 1024;  // grid of k_block_dot (four four-wave workgroups per CU)
// The update's byte counter is spread over this many slots (slot = workgroup index mod kByteSlots), so
// that the waves of a launch do not queue on one address; the host sums them.
constexpr int kByteSlots = 64;
__device__ __forceinline__ void add_bytes(unsigned long long *slots, uint64_t n) {
    atomicAdd(slots + (blockIdx.x & (kByteSlots - 1)), static_cast<unsigned long long>(n));
}
constexpr int kMaxLevels = 8;    // kChunk * 16^(kMaxLevels-1) covers 2^32 ranks

// Bases merged per group: one wave inserts a group's bases one after the other, so a group's time
// grows with its size, and each level is one launch.  Of 16 and 64, the size that gives fewer levels
// (16 on a tie).
inline int merge_group_size(int64_t chunks, int want) {
    if (want > 0) return want;
    auto levels = [&](int64_t gsz) {
        int l = 0;
        for (int64_t n = chunks; n > 1; n = (n + gsz - 1) / gsz) ++l;
        return l;
    };
    return levels(64) < levels(16) ? 64 : 16;
}
constexpr int kTile = 16;        // support words per workgroup of the XOR update (table form)
constexpr int kTileDirect = 32;  // the same, direct form (batched: many small components per step)

// One entry of the tree pass's work lists: a root's tree (its first position, its position count with
// bit 31 set for a narrow tree) and a rank-ascending run [cb, ce) of its candidates in the rc_* arrays.
// The lists hold these records in the order the workgroups and waves take them, so that an item costs
// one load, issued while the item before it runs.
struct ItemMeta {
    int64_t off, cb, ce;
    uint32_t npos, root;
};
constexpr uint32_t kNarrowBit = 0x80000000u;

struct BlockArgs {
    const int64_t *tree_off;
    const uint32_t *parent_pos;
    const int32_t *pos_coord;
    const uint32_t *pos_pre, *pos_end;
    const ItemMeta *big_list, *small_list;  // the work lists (big items narrow-flagged)
    const int32_t *wg_big, *wave_small;     // each workgroup's big items, each wave's small items
    uint64_t *scratch;          // [blocks][huge_words] path words of trees too big for LDS
    int64_t huge_words;
    const int64_t *rc_off;
    const uint32_t *rc_pu, *rc_pv, *rc_rank;
    const int32_t *rc_coord;
    const uint32_t *cand_root, *cand_pu, *cand_pv;
    const int32_t *cand_coord;
    uint64_t *Sw;
    int64_t npad;
    uint64_t *Tt;        // [npad] bit l of Tt[c] = bit c of T_l
    uint64_t *Qr;        // [K]    bit l of Qr[k] = <C_k, T_l>, valid where qbits is set
    uint64_t *qbits;     // [ceil(K/64)] ranks whose word was emitted by the tree pass
    uint32_t *sum_cnt;   // [slots]      greedy bases: size
    uint32_t *sum_rank;  // [slots][64]  ranks, ascending
    uint64_t *sum_q;     // [slots][64]  their words
    int64_t lvl_slot[kMaxLevels];  // first slot of each level
    int64_t lvl_n[kMaxLevels];     // bases on each level
    int32_t levels;
    uint32_t *arrive;    // [slots] members of each merge group done so far, by the group's slot (zero between launches)
    uint32_t *sel;       // [N] rank chosen in each phase
    uint64_t *blk_q;     // [64] word of phase i's winner
    uint64_t *blk_m;     // [64] mask m_i
    uint64_t *blk_g;     // [64] bit l of row i = <C_{e+i}, S_{e+l}>, l < i
    uint64_t *C_blk;     // [64][W] cycle vectors of the block
    uint32_t *D_blk;     // [64][W] nonzero words of each below e
    uint32_t *Dcnt_blk;  // [64]
    uint64_t *Sf_blk;    // [W][64] final support vectors of the block (word-major)
    uint64_t *C_store;   // optional [N][W]
    uint64_t *Ct;        // [npad] bit i of Ct[c] = bit c of C_{e+i}
    uint64_t *Bc;        // [npad] update coefficients of each later column
    uint32_t *ulist;     // [npad] the later columns whose coefficients are nonzero (any order)
    uint32_t *ucount;    // [1] their number (cleared by k_block_transpose)
    uint32_t *pw_list;   // [W] words below e where some C_{e+i} is nonzero
    uint32_t *pw_off;    // [W+1] first pair of each listed word (prefix sums)
    uint32_t *pw_cnt;    // [1] listed words
    uint32_t *pair_i;    // [64*W] phase i of each (word, C_{e+i}) pair
    uint64_t *pair_c;    // [64*W] that word of C_{e+i}
    uint32_t *status;
    unsigned long long *upd_bytes;  // accumulated algorithmic bytes of the update
    int64_t K;
    int32_t e, kk, N, W, n_big, n_small, big_words, small_words, group;
    int32_t chunk_words;  // bitmap words per first-level basis
    int32_t seg_in_order;  // narrow scratch trees: resolve every segment in order
};

// One block of kk phases of one problem in that problem's own numbering: its support matrix
// Sw [W][npad], its phase-indexed selection and the block's work arrays.  Tt, Ct and Bc start at
// the problem's first coordinate, which is coordinate cb of the tree and candidate arrays (0 for a
// single problem; the start of the problem's range in a batch).
struct BlkView {
    uint64_t *Sw;
    int64_t npad, cb;
    int32_t e, kk, N, W;
    uint32_t *sel;
    uint64_t *blk_q, *blk_m, *blk_g;
    uint64_t *C_blk, *Sf_blk;
    uint32_t *D_blk, *Dcnt_blk;
    uint64_t *C_store;
    uint64_t *Tt, *Ct, *Bc;
    uint32_t *pw_list, *pw_off, *pw_cnt, *pair_i;
    uint64_t *pair_c;
    uint32_t *ulist, *ucount;  // later columns that take an update (null: not listed)
};

__device__ __forceinline__ BlkView view_of(const BlockArgs &a) {
    BlkView v;
    v.Sw = a.Sw;
    v.npad = a.npad;
    v.cb = 0;
    v.e = a.e;
    v.kk = a.kk;
    v.N = a.N;
    v.W = a.W;
    v.sel = a.sel;
    v.blk_q = a.blk_q;
    v.blk_m = a.blk_m;
    v.blk_g = a.blk_g;
    v.C_blk = a.C_blk;
    v.Sf_blk = a.Sf_blk;
    v.D_blk = a.D_blk;
    v.Dcnt_blk = a.Dcnt_blk;
    v.C_store = a.C_store;
    v.Tt = a.Tt;
    v.Ct = a.Ct;
    v.Bc = a.Bc;
    v.pw_list = a.pw_list;
    v.pw_off = a.pw_off;
    v.pw_cnt = a.pw_cnt;
    v.pair_i = a.pair_i;
    v.pair_c = a.pair_c;
    v.ulist = a.ulist;
    v.ucount = a.ucount;
    return v;
}

// 64 x 64 bit transpose across a wave: lane r holds row r (bit c: entry (r, c)); lane c receives
// column c (bit r: entry (r, c)).  Six rounds of swapping the off-diagonal j x j blocks of every 2j x 2j
// block with the partner lane l ^ j, j = 32, 16, ..., 1.
__device__ __forceinline__ uint64_t transpose64(uint64_t x, int lane) {
    constexpr uint64_t kLow[6] = {0x00000000FFFFFFFFull, 0x0000FFFF0000FFFFull, 0x00FF00FF00FF00FFull,
                                  0x0F0F0F0F0F0F0F0Full, 0x3333333333333333ull, 0x5555555555555555ull};
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const int j = 32 >> k;
        const uint64_t m = kLow[k];
        const uint32_t lo = static_cast<uint32_t>(__shfl_xor(static_cast<int>(static_cast<uint32_t>(x)), j));
        const uint32_t hi = static_cast<uint32_t>(__shfl_xor(static_cast<int>(static_cast<uint32_t>(x >> 32)), j));
        const uint64_t y = (static_cast<uint64_t>(hi) << 32) | lo;
        x = (lane & j) ? ((x & ~m) | ((y & ~m) >> j)) : ((x & m) | ((y & m) << j));
    }
    return x;
}

// Transposes the block's vectors: wave w builds Tt[64w .. 64w+63] from word w of columns
// e .. e+kk-1 (bit b of that word in lane l is bit l of Tt[64w+b]); own bits are added.
__device__ __forceinline__ void transpose_body(const BlkView &v, int64_t w, int lane) {
    const uint64_t x = lane < v.kk ? v.Sw[w * v.npad + v.e + lane] : 0ull;
    uint64_t mine = transpose64(x, lane);
    const int64_t c = w * kWave + lane;
    if (c >= v.e && c < static_cast<int64_t>(v.e) + v.kk) mine |= 1ull << (c - v.e);
    v.Tt[c] = mine;
}

__global__ __launch_bounds__(256) void k_block_transpose(BlockArgs a) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *a.ucount = 0;  // k_block_dot lists this block's updated columns
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t w = static_cast<int64_t>(blockIdx.x) * 4 + wave_id();
    const int64_t wl = (static_cast<int64_t>(a.e) + a.kk + kWave - 1) / kWave;
    if (w >= wl) return;
    transpose_body(view_of(a), w, lane);
}

// The lowest set bit of x alone (0 for x = 0).
__device__ __forceinline__ uint64_t lowbit(uint64_t x) { return x & (0ull - x); }

// Value of a 64-bit register in lane l (l wave-uniform).
__device__ __forceinline__ uint64_t rl64(uint64_t x, int l) {
    const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(x)), l));
    const uint32_t hi =
        static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(x >> 32)), l));
    return (static_cast<uint64_t>(hi) << 32) | lo;
}

// Settles the lanes of one window whose parent lies in the same window, by pointer jumping: a
// pending lane holds the XOR of the terms on its path up to (not including) its in-window ancestor
// anc; each round it takes anc's value and anc's ancestor (both as they were before the round), so
// the distance to a settled ancestor halves and a window of 64 settles in at most six rounds.  A
// parent precedes its child, so every chain ends at a lane whose parent lies before the window.
__device__ __forceinline__ uint64_t settle_words(uint64_t val, bool pending, uint32_t p, uint32_t wb, int lane) {
    int anc = pending ? static_cast<int>(p - wb) : -1;
    while (__ballot(anc >= 0)) {
        const int src = anc >= 0 ? anc : lane;
        const uint64_t av = __shfl(val, src);
        const int aa = __shfl(anc, src);
        if (anc >= 0) {
            val ^= av;
            anc = aa;
        }
    }
    return val;
}

// Relative words of kU windows at once: in window u, a lane whose parent lies in the window starts with
// jmp[u] = that parent's lane, the others with jmp[u] = -1 and anc[u] = their parent (before the
// window, or kNoParent).  Pointer jumping, all windows in lockstep so that their lane permutes are in
// flight together: a pending lane takes its in-window ancestor's value and ancestor (both as they were
// before the round), and on landing on a settled lane, that lane's anc.  Afterwards rel[u] is the XOR
// of the terms on the lane's path up to its first ancestor before the window (excluded) and anc[u] that
// ancestor.  A parent precedes its child, so every chain ends within six rounds.
template <int kU>
__device__ __forceinline__ void settle_windows(uint64_t (&rel)[kU], uint32_t (&anc)[kU], int (&jmp)[kU], int lane) {
    for (;;) {
        bool pend = false;
#pragma unroll
        for (int u = 0; u < kU; ++u) pend |= jmp[u] >= 0;
        if (!__ballot(pend)) break;
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int src = jmp[u] >= 0 ? jmp[u] : lane;
            const uint64_t av = __shfl(rel[u], src);
            const int aj = __shfl(jmp[u], src);
            const uint32_t aa = static_cast<uint32_t>(__shfl(static_cast<int>(anc[u]), src));
            if (jmp[u] >= 0) {
                rel[u] ^= av;
                jmp[u] = aj;
                if (aj < 0) anc[u] = aa;
            }
        }
    }
}

// Path words of U windows of one wave's tree: word[pos] = word[parent] ^ Tt[coordinate], t[u] being
// window u's Tt words (gathered while the group before ran).  kRing: Wd is global scratch and ring
// (rmask + 1 words of LDS, a power of two of at least two windows) keeps the most recent windows'
// words, so a parent in one of them is read from LDS.
template <bool kRing, int U>
__device__ __forceinline__ void word_group(const WindowGroupT<U> &g, const uint64_t (&t)[U], uint32_t base,
                                           uint32_t npos, uint64_t *Wd, uint64_t *ring, uint32_t rmask, int lane) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const uint32_t wb = base + u * kWave;
        if (wb >= npos) break;
        const bool v = wb + lane < npos;
        const uint32_t p = v ? g.par[u] : kNoParent;
        const uint64_t tv = (v && g.co[u] >= 0) ? t[u] : 0ull;
        const bool before = p < wb;
        uint64_t pw = 0;
        if (before) {
            // the ring holds positions [wb + 64 - ring size, wb) while window wb is walked
            if (kRing && p + rmask + 1 >= wb + kWave) pw = ring[p & rmask];
            else pw = Wd[p];
        }
        uint64_t val = tv ^ pw;
        const bool inwin = (p != kNoParent) && (p >= wb);
        if (__ballot(inwin)) val = settle_words(val, inwin, p, wb, lane);
        if (v) {
            Wd[wb + lane] = val;
            if (kRing) ring[(wb + lane) & rmask] = val;
        }
    }
}

template <int U>
__device__ __forceinline__ void gather_windows(uint64_t (&t)[U], const WindowGroupT<U> &g, const uint64_t *Tt) {
#pragma unroll
    for (int u = 0; u < U; ++u) t[u] = Tt[g.co[u] >= 0 ? g.co[u] : 0];
}

// One wave's walk of a tree, kTwUnroll windows per group and three groups in flight: the current
// group's words, the next group's Tt words, the positions of the group after.
template <bool kRing = false>
__device__ __forceinline__ void tree_words(const uint32_t *pp, const int32_t *pc, uint32_t npos, const uint64_t *Tt,
                                           uint64_t *Wd, int lane, uint64_t *ring = nullptr, uint32_t rmask = 0) {
    constexpr int U = kTwUnroll;
    constexpr uint32_t G = kWave * U;
    WindowGroupT<U> cur, nxt;
    uint64_t tc[U], tn[U];
    load_group(cur, pp, pc, npos, lane, kWave);
    gather_windows(tc, cur, Tt);
    load_group(nxt, pp, pc, npos, G + lane, kWave);
    for (uint32_t base = 0; base < npos; base += G) {
        gather_windows(tn, nxt, Tt);
        WindowGroupT<U> later;
        load_group(later, pp, pc, npos, base + 2 * G + lane, kWave);
        word_group<kRing, U>(cur, tc, base, npos, Wd, ring, rmask, lane);
        cur = nxt;
        nxt = later;
#pragma unroll
        for (int u = 0; u < U; ++u) tc[u] = tn[u];
    }
}

// One group of a segment walk (positions [s0, s1) of a narrow tree, one wave): Wd[q] becomes the XOR of
// the Tt words on the path from q up to its first ancestor before s0, excluded, and anch[q] that
// ancestor (kNoParent when the path reaches the root inside the segment).  rv/ra: the wave's ring of its
// latest rmask + 1 positions (words and anchors) in LDS.  The group's windows settle together
// (settle_windows) to words relative to each window's entry; only the last step, joining a window to
// its ancestors' words in the ring, runs window after window, so the serial chain of the walk is one
// LDS read and write per window.
constexpr int kSegUnroll = 2;  // windows per prefetched group of a segment walk (settled in lockstep)
using SegWindows = WindowGroupT<kSegUnroll>;

// A group's words and anchors for scratch, stored at the start of the next group: after the loads that the next
// wait is for, so that the wait (the vector memory counter retires in order) does not wait on these stores.
struct SegPending {
    uint64_t val[kSegUnroll];
    uint32_t an[kSegUnroll];
    uint32_t base;  // kNoParent: nothing pending
};

__device__ __forceinline__ void flush_pending(const SegPending &pd, uint32_t s1, uint64_t *Wd, uint32_t *anch, int lane) {
    if (pd.base == kNoParent) return;
#pragma unroll
    for (int u = 0; u < kSegUnroll; ++u) {
        const uint32_t q = pd.base + u * kWave + lane;
        if (q < s1) {
            Wd[q] = pd.val[u];
            anch[q] = pd.an[u];
        }
    }
}

__device__ __forceinline__ void segment_group(const SegWindows &g, const uint64_t (&t)[kSegUnroll], uint32_t base,
                                              uint32_t s0, uint32_t s1, uint64_t *Wd, uint32_t *anch, uint64_t *rv,
                                              uint32_t *ra, uint32_t rmask, uint32_t near_lo, uint32_t seg0_end,
                                              bool &far_anchor, SegPending &pd, int lane) {
    flush_pending(pd, s1, Wd, anch, lane);
    uint64_t rel[kSegUnroll];
    uint32_t anc[kSegUnroll];
    int jmp[kSegUnroll];
#pragma unroll
    for (int u = 0; u < kSegUnroll; ++u) {
        const uint32_t wb = base + u * kWave;
        const bool v = wb + lane < s1;
        const uint32_t p = v ? g.par[u] : kNoParent;
        rel[u] = (v && g.co[u] >= 0) ? t[u] : 0ull;
        const bool inwin = p != kNoParent && p >= s0 && p >= wb;
        jmp[u] = inwin ? static_cast<int>(p - wb) : -1;
        anc[u] = inwin ? kNoParent : p;
    }
    settle_windows<kSegUnroll>(rel, anc, jmp, lane);
#pragma unroll
    for (int u = 0; u < kSegUnroll; ++u) {
        const uint32_t wb = base + u * kWave;
        if (wb >= s1) break;
        const bool v = wb + lane < s1;
        // an ancestor inside the segment (before this window) contributes its relative word and passes
        // on its segment anchor; one before the segment (or kNoParent) is the segment anchor itself
        const uint32_t x = v ? anc[u] : kNoParent;
        const bool inseg = x != kNoParent && x >= s0;
        const bool near = inseg && x + rmask + 1 >= wb + kWave;
        const bool far = inseg && !near;
        // an ancestor older than the ring: this wave's own earlier stores must be complete before it is read
        if (__ballot(far)) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        }
        uint64_t pw = 0;
        uint32_t an = x;
        if (near) {
            pw = rv[x & rmask];
            an = ra[x & rmask];
        } else if (far) {
            pw = Wd[x];
            an = anch[x];
        }
        const uint64_t val = rel[u] ^ pw;
        pd.val[u] = val;
        pd.an[u] = an;
        if (v) {
            const uint32_t q = wb + lane;
            rv[q & rmask] = val;
            ra[q & rmask] = an;
            // a segment anchor outside segment 0 and before the previous segment's ring
            far_anchor |= an != kNoParent && an >= seg0_end && an < near_lo;
        }
    }
    pd.base = base;
}

// A narrow tree in scratch walked by all nw waves of the workgroup: wave s walks the positions
// [s·seg, (s+1)·seg) relative to the segment (segment_group).  When every segment's anchors lie in segment 0
// or in the previous segment's last T = rmask + 1 positions (still in that segment's LDS ring), only the rings
// are resolved, segment after segment, in LDS (their final words and kNoParent anchors go to scratch); every
// other position keeps its relative word and anchor, whose word is now final, and the candidate run joins the
// two at the ends it reads (run_words with anch): returns true.  Otherwise the segments after the first are
// resolved in order, every thread taking positions of one segment at a time (the word of q is its relative
// word XOR the final word of its anchor), and it returns false.
__device__ bool tree_words_segmented(const uint32_t *pp, const int32_t *pc, uint32_t npos, const uint64_t *Tt,
                                     uint64_t *Wd, uint32_t *anch, uint64_t *ring_base, uint32_t rmask, int wid,
                                     int nw, bool in_order, int lane) {
    const uint32_t seg = ((npos + nw - 1) / nw + kWave - 1) / kWave * kWave;
    const uint32_t s0 = static_cast<uint32_t>(wid) * seg, s1 = min(npos, s0 + seg);
    const uint32_t T = rmask + 1;
    uint64_t *rv = ring_base + static_cast<size_t>(wid) * T;
    uint32_t *ra = reinterpret_cast<uint32_t *>(ring_base + static_cast<size_t>(nw) * T) + static_cast<size_t>(wid) * T;
    bool far_anchor = false;
    if (s0 < s1) {
        // three groups in flight: this one's words, the next one's Tt words, the one after's positions
        constexpr uint32_t G = kWave * kSegUnroll;
        SegWindows cur, nxt;
        uint64_t tc[kSegUnroll], tn[kSegUnroll];
        load_group(cur, pp, pc, s1, s0 + lane, kWave);
        gather_windows(tc, cur, Tt);
        load_group(nxt, pp, pc, s1, s0 + G + lane, kWave);
        const uint32_t near_lo = s0 > T ? s0 - T : 0u;
        SegPending pd;
        pd.base = kNoParent;
        for (uint32_t base = s0; base < s1; base += G) {
            gather_windows(tn, nxt, Tt);
            SegWindows later;
            load_group(later, pp, pc, s1, base + 2 * G + lane, kWave);
            segment_group(cur, tc, base, s0, s1, Wd, anch, rv, ra, rmask, near_lo, seg, far_anchor, pd, lane);
            cur = nxt;
            nxt = later;
#pragma unroll
            for (int u = 0; u < kSegUnroll; ++u) tc[u] = tn[u];
        }
        flush_pending(pd, s1, Wd, anch, lane);
    }
    if (!__syncthreads_or(far_anchor || seg < T || in_order)) {
        for (uint32_t sg = 1; sg < static_cast<uint32_t>(nw); ++sg) {
            const uint32_t a0 = sg * seg;
            if (a0 >= npos) break;
            const uint32_t a1 = min(npos, a0 + seg);
            const uint32_t t0 = a1 - a0 > T ? a1 - T : a0;
            uint64_t *rw = ring_base + static_cast<size_t>(sg) * T;
            uint32_t *rn = reinterpret_cast<uint32_t *>(ring_base + static_cast<size_t>(nw) * T) + static_cast<size_t>(sg) * T;
            const uint64_t *pw = ring_base + static_cast<size_t>(sg - 1) * T;  // the previous segment's ring, final
            for (uint32_t q = t0 + threadIdx.x; q < a1; q += blockDim.x) {
                const uint32_t x = rn[q & rmask];
                uint64_t w = rw[q & rmask];
                if (x != kNoParent) w ^= x >= seg ? pw[x & rmask] : Wd[x];
                rw[q & rmask] = w;
                Wd[q] = w;
                anch[q] = kNoParent;
            }
            __syncthreads();
        }
        return true;
    }
    for (int sg = 1; sg < nw; ++sg) {
        const uint32_t a0 = static_cast<uint32_t>(sg) * seg;
        if (a0 >= npos) break;
        const uint32_t a1 = min(npos, a0 + seg);
        constexpr int kR = 4;
        for (uint32_t q0 = a0 + threadIdx.x; q0 < a1; q0 += kR * blockDim.x) {
            uint32_t an[kR];
            uint64_t w[kR], aw[kR];
#pragma unroll
            for (int k = 0; k < kR; ++k) {
                const uint32_t q = q0 + k * blockDim.x;
                an[k] = anch[q < a1 ? q : a0];
            }
#pragma unroll
            for (int k = 0; k < kR; ++k) {
                const uint32_t q = q0 + k * blockDim.x;
                w[k] = Wd[q < a1 ? q : a0];
                aw[k] = Wd[an[k] != kNoParent ? an[k] : 0u];
            }
#pragma unroll
            for (int k = 0; k < kR; ++k) {
                const uint32_t q = q0 + k * blockDim.x;
                if (q < a1 && an[k] != kNoParent) Wd[q] = w[k] ^ aw[k];
            }
        }
        __syncthreads();
    }
    return false;
}

// The same for a big tree shared by all waves of the workgroup: window w belongs to wave w mod nw
// and is published with a ready stamp (this root's generation) once its words are in LDS.  A wave
// waits only for lower windows, and the lowest unfinished window never waits, so every wait ends.
__device__ __forceinline__ void coop_word_group(const WindowGroup &g, uint32_t w0, uint32_t nwin, uint32_t npos,
                                                const uint64_t *Tt, uint64_t *Wd, uint32_t *ready, uint32_t gen,
                                                int nw, int lane) {
    uint64_t t[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) t[u] = Tt[g.co[u] >= 0 ? g.co[u] : 0];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
        const uint32_t w = w0 + u * nw;
        if (w >= nwin) break;
        const uint32_t wb = w * kWave;
        const bool v = wb + lane < npos;
        const uint32_t p = v ? g.par[u] : kNoParent;
        const uint64_t tv = (v && g.co[u] >= 0) ? t[u] : 0ull;
        const bool before = p < wb;
        const uint32_t pw = before ? p / kWave : 0u;
        while (before && __hip_atomic_load(ready + pw, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) != gen)
            __builtin_amdgcn_s_sleep(1);
        uint64_t val = tv ^ (before ? Wd[p] : 0ull);
        const bool inwin = (p != kNoParent) && (p >= wb);
        if (__ballot(inwin)) val = settle_words(val, inwin, p, wb, lane);
        if (v) Wd[wb + lane] = val;
        if (lane == 0) __hip_atomic_store(ready + w, gen, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
}

__device__ __forceinline__ void tree_words_coop(const uint32_t *pp, const int32_t *pc, uint32_t npos,
                                                const uint64_t *Tt, uint64_t *Wd, uint32_t *ready, uint32_t gen,
                                                int wid, int nw, int lane) {
    const uint32_t nwin = (npos + kWave - 1) / kWave;
    const uint32_t stride = static_cast<uint32_t>(nw) * kWave;
    const uint32_t gstep = kUnroll * static_cast<uint32_t>(nw);
    WindowGroup A, B;
    load_group(A, pp, pc, npos, wid * kWave + lane, stride);
    for (uint32_t w0 = wid; w0 < nwin; w0 += 2 * gstep) {
        load_group(B, pp, pc, npos, (w0 + gstep) * kWave + lane, stride);
        coop_word_group(A, w0, nwin, npos, Tt, Wd, ready, gen, nw, lane);
        if (w0 + gstep >= nwin) break;
        load_group(A, pp, pc, npos, (w0 + 2 * gstep) * kWave + lane, stride);
        coop_word_group(B, w0 + gstep, nwin, npos, Tt, Wd, ready, gen, nw, lane);
    }
}

// New pivots from a batch already reduced by the pivots (lane i: word v), one per step, without
// branches inside the step: the lowest nonzero lane p gives its word, whose lowest bit c is a new
// pivot column; lane c's pivot becomes that word, and every other pivot and every lane whose word
// has bit c take it (so every pivot stays zero in every other pivot column, and lane p's word becomes
// zero).  The chain from one ballot to the next is a lane read, two bit scans and masked XORs.
// Returns whether this lane became a pivot; added counts the new pivots (wave-uniform).
__device__ __forceinline__ bool pivot_steps(uint64_t &piv, uint64_t &P, uint64_t v, int lane, int &added) {
    bool mine = false;
    uint64_t nz = __ballot(v != 0);
    while (nz) {
        const int p = __builtin_ctzll(nz);
        const uint64_t pvv = rl64(v, p);  // nonzero: lane p's word
        const int c = __builtin_ctzll(pvv);
        mine = mine || lane == p;
        const bool vb = ((v >> c) & 1ull) != 0;
        const bool pb = ((piv >> c) & 1ull) != 0;
        piv = lane == c ? pvv : (pb ? piv ^ pvv : piv);
        P |= 1ull << c;
        ++added;
        v ^= vb ? pvv : 0ull;
        nz = __ballot(v != 0);
    }
    return mine;
}

// Greedy insertion of one batch (lane i: word q, ranks ascending with the lane; q = 0 for no entry)
// into a wave basis kept in reduced echelon form: lane b holds the pivot of column b (P: the pivot
// columns), and every pivot is zero in every other pivot column.  The batch is reduced by the
// pivots (one XOR per pivot column the word has); then, while a lane is nonzero, the first such
// lane becomes the pivot of its lowest set bit, that bit is cleared from the other pivots and from
// the later lanes.  The lanes that become pivots are exactly the batch vectors independent of the
// basis and of the batch vectors before them; emit() is called in each of them (together, once the
// batch is inserted).  Returns the number of new pivots (wave-uniform).
template <typename EmitFn>
__device__ __forceinline__ int basis_reduce_insert(uint64_t &piv, uint64_t &P, uint64_t q, int lane, EmitFn emit) {
    uint64_t v = q;
    P = static_cast<uint64_t>(uniform64(static_cast<int64_t>(P)));
    if (__ballot((v & P) != 0)) {
        uint64_t r = 0;
        for (uint64_t pm = P; pm; pm &= pm - 1) {
            const int c = __builtin_ctzll(pm);
            const uint64_t pc = rl64(piv, c);
            r ^= ((v >> c) & 1ull) ? pc : 0ull;
        }
        v ^= r;
    }
    // A lane becomes a pivot at most once (its own word is cleared then), so its emit() can wait
    // until the loop ends.
    int added = 0;
    const bool mine = pivot_steps(piv, P, v, lane, added);
    P = static_cast<uint64_t>(uniform64(static_cast<int64_t>(P)));
    if (mine) emit();
    return added;
}

// Candidate words of one rank-ascending run of a root's candidates, kept only where they are
// independent of the run's earlier words: word = path word[u] ^ path word[v] ^ Tt[own coordinate].
// The wave keeps the run's greedy basis (lane b: pivot with lowest set bit b; P: its columns).  A
// batch is first reduced by the existing pivots lane by lane; the lanes left nonzero are resolved
// column by column, the first lane with a column's bit becoming its pivot.  Each new pivot's rank
// and original word go to Qr/qbits.  A word dependent on earlier words of the same run is dependent
// on earlier candidates overall, so it cannot be in the block's greedy basis and is dropped.
template <typename A>
__device__ __forceinline__ void run_words(const A &a, const uint64_t *Wd, int64_t k_begin, int64_t k_end, int kk,
                                          int lane, const uint32_t *anch = nullptr) {
    constexpr int kB = 4;
    uint64_t piv = 0, P = 0;
    int npiv = 0;
    for (int64_t k0 = k_begin; k0 < k_end && npiv < kk; k0 += kB * kWave) {
        uint32_t pu[kB], pv[kB], rk[kB];
        int32_t c[kB];
#pragma unroll
        for (int b = 0; b < kB; ++b) {
            const int64_t k = k0 + b * kWave + lane;
            const int64_t at = k < k_end ? k : k_begin;
            pu[b] = a.rc_pu[at];
            pv[b] = a.rc_pv[at];
            rk[b] = a.rc_rank[at];
            c[b] = a.rc_coord[at];
        }
        // every batch's words first, then the insertions and their emissions: the vector memory counter
        // retires in order, so a load issued after an emission's store would wait for that store
        uint64_t qb[kB];
#pragma unroll
        for (int b = 0; b < kB; ++b) {
            const int64_t k = k0 + b * kWave + lane;
            uint64_t q = 0;
            if (k < k_end) {
                q = Wd[pu[b]] ^ Wd[pv[b]];
                if (anch) {  // words relative to a segment's entry: XOR in the final word of each end's anchor
                    const uint32_t xu = anch[pu[b]], xv = anch[pv[b]];
                    if (xu != kNoParent) q ^= Wd[xu];
                    if (xv != kNoParent) q ^= Wd[xv];
                }
                if (c[b] >= 0) q ^= a.Tt[c[b]];
            }
            qb[b] = q;
        }
#pragma unroll
        for (int b = 0; b < kB; ++b) {
            const uint64_t q = qb[b];
            npiv += basis_reduce_insert(piv, P, q, lane, [&]() {
                a.Qr[rk[b]] = q;
                atomicOr(reinterpret_cast<unsigned long long *>(a.qbits + (rk[b] >> 6)), 1ull << (rk[b] & 63));
            });
        }
    }
}

__device__ __forceinline__ bool seg_in_order_of(const BlockArgs &a) { return a.seg_in_order != 0; }
template <typename A>
__device__ __forceinline__ bool seg_in_order_of(const A &) { return false; }

// A small item of at most one window of positions and one batch of candidates.
__device__ __forceinline__ bool is_tiny(const ItemMeta &m) {
    return (m.npos & ~kNarrowBit) <= static_cast<uint32_t>(kWave) && m.ce - m.cb <= kWave;
}

// What a tiny item reads from its tree layout and candidate run: lane x its position x (parent,
// coordinate) and its candidate cb + x (ends, rank, own coordinate); loads of a lane past the item's
// end are clamped to its first entry, and nothing is loaded for a non-tiny or absent item.
struct TinyLoads {
    uint32_t par, pu, pv, rk;
    int32_t co, c;
};

template <typename A>
__device__ __forceinline__ TinyLoads load_tiny(const A &a, const ItemMeta &m, bool present, int lane) {
    TinyLoads t{kNoParent, 0u, 0u, 0u, -1, -1};
    if (present && is_tiny(m)) {
        const uint32_t npos = m.npos & ~kNarrowBit;
        const int64_t at = m.off + (static_cast<uint32_t>(lane) < npos ? lane : 0);
        t.par = a.parent_pos[at];
        t.co = a.pos_coord[at];
        const int64_t k = m.cb + lane < m.ce ? m.cb + lane : m.cb;
        t.pu = a.rc_pu[k];
        t.pv = a.rc_pv[k];
        t.rk = a.rc_rank[k];
        t.c = a.rc_coord[k];
    }
    return t;
}

// A tiny item from its loads: the path words of its one window settle by pointer jumping (every
// parent lies in the window), each candidate's word is its ends' words (lane permutes) ^ Tt[own
// coordinate], and the batch goes through the run's greedy basis as run_words would insert it.
template <typename A>
__device__ __forceinline__ void tiny_item(const A &a, const TinyLoads &t, uint32_t npos, uint32_t ncand, int lane) {
    const bool v = static_cast<uint32_t>(lane) < npos;
    const bool kv = static_cast<uint32_t>(lane) < ncand;
    const uint64_t tv = (v && t.co >= 0) ? a.Tt[t.co] : 0ull;
    const uint64_t tc = (kv && t.c >= 0) ? a.Tt[t.c] : 0ull;
    const uint32_t p = v ? t.par : kNoParent;
    const bool inwin = p != kNoParent;
    uint64_t val = tv;
    if (__ballot(inwin)) val = settle_words(val, inwin, p, 0u, lane);
    const uint64_t wu = __shfl(val, static_cast<int>(t.pu & (kWave - 1)));
    const uint64_t wv = __shfl(val, static_cast<int>(t.pv & (kWave - 1)));
    const uint64_t q = kv ? (wu ^ wv ^ tc) : 0ull;
    uint64_t piv = 0, P = 0;
    const uint32_t rk = t.rk;
    basis_reduce_insert(piv, P, q, lane, [&]() {
        a.Qr[rk] = q;
        atomicOr(reinterpret_cast<unsigned long long *>(a.qbits + (rk >> 6)), 1ull << (rk & 63));
    });
}

// Pass over every tree and every candidate of the block.  The work lists hold items (a root and a
// rank-ascending run of its candidates, as ItemMeta records): big trees (more than
// small_words positions) take a whole workgroup per item, the others one wave per item.  This
// workgroup takes the big items big[first], big[first + stride], ... below big[end] (narrow[i]: item
// i's tree is narrow) and this wave the small items small[sfirst], small[sfirst + sstride], ... below
// small[send].
// A big tree of more than big_words positions keeps its path words in this workgroup's region of
// scratch (huge_words words) instead of LDS; its ready stamps stay in LDS and are cleared first, so
// that no stale word can pass for a stamp.  kk_of(r) is the number of phases in root r's block.
// Work-list entry i (clamped below end, so that the load is unconditional; zeros for an empty list),
// wave-uniform.
__device__ __forceinline__ ItemMeta load_item(const ItemMeta *list, uint32_t i, uint32_t end) {
    ItemMeta m{};
    if (end == 0) return m;
    const ItemMeta *p = list + (i < end ? i : end - 1);
    m.off = uniform64(p->off);
    m.cb = uniform64(p->cb);
    m.ce = uniform64(p->ce);
    m.npos = uniform32(p->npos);
    m.root = uniform32(p->root);
    return m;
}

template <typename A, typename KkFn>
__device__ __forceinline__ void tree_pass(const A &a, const ItemMeta *big, uint32_t first, uint32_t end,
                                          uint32_t stride, const ItemMeta *small, uint32_t sfirst, uint32_t send,
                                          uint32_t sstride, KkFn kk_of) {
    extern __shared__ uint64_t lds[];
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = wave_id();
    const int nw = blockDim.x / kWave;
    uint64_t *Wbig = lds;
    uint32_t *ready = reinterpret_cast<uint32_t *>(lds + a.big_words);
    const uint32_t nwin_big = static_cast<uint32_t>(a.big_words) / kWave;
    // ring of a scratch tree's latest windows: the largest power of two of words in the LDS tree
    // region, if it holds at least two windows (the ready stamps of a narrow tree are not used)
    uint32_t ring_mask = 0;
    if (a.big_words >= 2 * kWave) ring_mask = (1u << (31 - __builtin_clz(static_cast<uint32_t>(a.big_words)))) - 1;
    // per-wave rings of a segmented walk (12 bytes per position, words then anchors) in the same LDS tree
    // region: the largest power of two of positions that fits nw times, if it holds two windows
    uint32_t seg_mask = 0;
    if (nw >= 2) {
        const uint32_t per = static_cast<uint32_t>(a.big_words) * 8u / static_cast<uint32_t>(nw) / 12u;
        if (per >= 2 * kWave) seg_mask = (1u << (31 - __builtin_clz(per))) - 1;
    }
    for (uint32_t i = threadIdx.x; i < nwin_big; i += blockDim.x) ready[i] = 0;
    __syncthreads();
    uint32_t gen = 0;
    ItemMeta next = load_item(big, first, end);
    for (uint32_t it = first; it < end; it += stride) {
        const ItemMeta cur = next;
        next = load_item(big, it + stride, end);  // in flight while this item runs
        const uint32_t r = cur.root;
        const int64_t off = cur.off;
        const uint32_t npos = cur.npos & ~kNarrowBit;
        ++gen;
        const bool nar = (cur.npos & kNarrowBit) != 0;
        const int64_t cb = cur.cb, ce = cur.ce;
        // path words, then each wave takes one contiguous run of the item's candidates (the two calls
        // keep the LDS and the scratch accesses in their own address spaces)
        auto one_tree = [&](uint64_t *Wd, uint32_t *rd, bool in_scratch) {
            uint32_t *seg_anch = nullptr;  // set while a segmented walk's words stay relative
            if (nar) {
                // narrow, deep tree: windows mostly wait on the window before them, so one wave walks
                // it alone (no handoffs between waves); the others wait at the barrier.  A tree in
                // scratch keeps its latest windows in the idle LDS tree region as well (ring), so
                // that the walk does not wait on global memory for every window.
                if (in_scratch && seg_mask) {
                    // all waves, a segment each, then the segments' rings resolved in order (the rest at the
                    // candidates' ends) or, failing that, every segment in order
                    seg_anch = reinterpret_cast<uint32_t *>(a.scratch + static_cast<int64_t>(gridDim.x) * a.huge_words) +
                               static_cast<int64_t>(blockIdx.x) * a.huge_words;
                    if (!tree_words_segmented(a.parent_pos + off, a.pos_coord + off, npos, a.Tt, Wd, seg_anch, Wbig,
                                              seg_mask, wid, nw, seg_in_order_of(a), lane))
                        seg_anch = nullptr;
                } else if (wid == 0) {
                    if (in_scratch && ring_mask) {
                        tree_words<true>(a.parent_pos + off, a.pos_coord + off, npos, a.Tt, Wd, lane, Wbig, ring_mask);
                    } else {
                        tree_words(a.parent_pos + off, a.pos_coord + off, npos, a.Tt, Wd, lane);
                    }
                }
            } else {
                tree_words_coop(a.parent_pos + off, a.pos_coord + off, npos, a.Tt, Wd, rd, gen, wid, nw, lane);
            }
            __syncthreads();
            const int64_t seg = (ce - cb + nw - 1) / nw;
            const int64_t s0 = cb + seg * wid, s1 = s0 + seg < ce ? s0 + seg : ce;
            if (s0 < s1) run_words(a, Wd, s0, s1, kk_of(r), lane, seg_anch);
            __syncthreads();
        };
        if (npos <= static_cast<uint32_t>(a.big_words)) {
            one_tree(Wbig, ready, false);
        } else {
            uint32_t *rd = reinterpret_cast<uint32_t *>(lds);
            for (uint32_t i = threadIdx.x; i < (npos + kWave - 1) / kWave; i += blockDim.x) rd[i] = 0;
            __syncthreads();
            one_tree(a.scratch + static_cast<int64_t>(blockIdx.x) * a.huge_words, rd, true);
        }
    }
    // the big-tree region is free from here on; each wave takes its own slice for small roots.
    // A tiny item (one window of positions, one batch of candidates) is software-pipelined: its
    // positions and candidates are loaded while the wave works on the item before it, and its path
    // words stay in registers (a candidate reads its ends' words by lane permutes).
    uint64_t *Ws = lds + static_cast<size_t>(wid) * a.small_words;
    ItemMeta cur = load_item(small, sfirst, send);
    ItemMeta nxt = load_item(small, sfirst + sstride, send);
    TinyLoads tl = load_tiny(a, cur, sfirst < send, lane);
    for (uint32_t it = sfirst; it < send; it += sstride) {
        const ItemMeta later = load_item(small, it + 2 * sstride, send);  // in flight two items ahead
        const TinyLoads tn = load_tiny(a, nxt, it + sstride < send, lane);  // in flight while this item runs
        const uint32_t npos = cur.npos & ~kNarrowBit;
        if (is_tiny(cur)) {
            if (kk_of(cur.root) > 0) tiny_item(a, tl, npos, static_cast<uint32_t>(cur.ce - cur.cb), lane);
        } else {
            tree_words(a.parent_pos + cur.off, a.pos_coord + cur.off, npos, a.Tt, Ws, lane);
            run_words(a, Ws, cur.cb, cur.ce, kk_of(cur.root), lane);
        }
        cur = nxt;
        nxt = later;
        tl = tn;
    }
}

// Seven waves per SIMD: the walks and runs mostly wait on memory, and the 72 VGPRs that allows (three spilled
// to scratch) cost less than the waves gain (Delaunay 100k tree pass 48.0 -> 47.8 us, ER 20 000 235.5 -> 233.7 us).
__global__ __launch_bounds__(kQWaves * kWave, 7) void k_block_q(BlockArgs a) {
    const uint32_t nw = blockDim.x / kWave;
    const uint32_t c = blockIdx.x * nw + static_cast<uint32_t>(wave_id());
    tree_pass(a, a.big_list, static_cast<uint32_t>(a.wg_big[blockIdx.x]),
              static_cast<uint32_t>(a.wg_big[blockIdx.x + 1]), 1u, a.small_list,
              static_cast<uint32_t>(a.wave_small[c]), static_cast<uint32_t>(a.wave_small[c + 1]), 1u,
              [&](uint32_t) { return a.kk; });
}

// Greedy insertion of one batch into a wave basis; the new independent vectors (original word and
// rank) are appended in rank order to (out_rank, out_q); cnt (wave-uniform) stops at kk.
__device__ __forceinline__ void basis_insert(uint64_t &piv, uint64_t &P, uint32_t &cnt, uint64_t q, uint32_t rank,
                                             uint32_t *out_rank, uint64_t *out_q, int lane) {
    uint64_t mine = 0;
    basis_reduce_insert(piv, P, q, lane, [&]() { mine = 1; });
    // each new pivot lane writes its own entry, after the new pivots of the lanes before it
    const uint64_t newp = __ballot(mine != 0);
    if (mine) {
        const uint32_t at = cnt + static_cast<uint32_t>(__popcll(newp & ((1ull << lane) - 1)));
        out_rank[at] = rank;
        out_q[at] = q;
    }
    cnt += static_cast<uint32_t>(__popcll(newp));
}

// basis_insert with the reduction by existing pivots done through tables in LDS: tab[g][x] is the
// XOR of the pivots of the columns 4g+t for the set bits t of x (zero for columns without a pivot),
// so a word reduces with sixteen independent lookups.  The tables are rebuilt, when pivots changed
// since they were built, before a batch that has bits in pivot columns: lane l builds entries
// 4(l%4) .. 4(l%4)+3 of group l/4 from that group's four pivots.  tab: kMergeTab words for this wave.
constexpr int kMergeTab = 16 * 16;
__device__ __forceinline__ void basis_insert_tab(uint64_t &piv, uint64_t &P, uint32_t &cnt, bool &stale,
                                                 uint64_t *tab, uint64_t q, uint32_t rank, uint32_t *out_rank,
                                                 uint64_t *out_q, int lane) {
    uint64_t v = q;
    if (__ballot((v & P) != 0)) {
        if (stale) {
            const int g = lane >> 2;
            uint64_t pt[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int src = 4 * g + t;
                const uint32_t lo = static_cast<uint32_t>(__shfl(static_cast<int>(static_cast<uint32_t>(piv)), src));
                const uint32_t hi = static_cast<uint32_t>(__shfl(static_cast<int>(static_cast<uint32_t>(piv >> 32)), src));
                pt[t] = (static_cast<uint64_t>(hi) << 32) | lo;
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int x = 4 * (lane & 3) + k;
                uint64_t e = 0;
#pragma unroll
                for (int t = 0; t < 4; ++t) e ^= pt[t] & (0ull - static_cast<uint64_t>((x >> t) & 1));
                tab[g * 16 + x] = e;
            }
            __builtin_amdgcn_wave_barrier();
            stale = false;
        }
        uint64_t r = 0;
#pragma unroll
        for (int g = 0; g < 16; ++g) r ^= tab[g * 16 + static_cast<int>((v >> (4 * g)) & 15ull)];
        v ^= r;
    }
    if (__ballot(v != 0)) stale = true;
    int added = 0;
    const bool mine = pivot_steps(piv, P, v, lane, added);
    // each new pivot lane writes its own entry, after the new pivots of the lanes before it
    const uint64_t newp = __ballot(mine != 0);
    if (mine) {
        const uint32_t at = cnt + static_cast<uint32_t>(__popcll(newp & ((1ull << lane) - 1)));
        out_rank[at] = rank;
        out_q[at] = q;
    }
    cnt += static_cast<uint32_t>(__popcll(newp));
}

// The block's greedy basis (slot top) in rank order -> phase of each winner, masks m_i, the
// products <C_{e+i}, S_{e+l}> and the chosen ranks (sel[i], i < kk).  One wave.
//
// Lane l holds column l of the rank-ordered word matrix (bit r: bit l of winner r's word, i.e. the
// product of winner r with T_l) and the mask m_l of phase l (initially T_l alone).  Phase b takes the
// first winner r not yet taken whose bit in column b is set; every later phase l whose column has a
// set bit at r (an odd product with the winner under its current mask) takes m_l ^= m_b and, the
// products being linear in the mask, column l ^= column b.  This is the sequential update of phases
// e..e+kk-1 on the block's words.  The winner chosen for each column is the one the row-wise
// elimination in rank order picks, since both read the same Schur complement entry.  At the end
// column l holds the products with the final S_{e+l}, so bit l of the products of phase i's winner
// is column l's bit at that winner: one ballot per phase.
__device__ void finalize_core(int kk, const uint32_t *sum_cnt, const uint32_t *sum_rank, const uint64_t *sum_q,
                              int64_t top, uint32_t *sel, uint64_t *blk_q, uint64_t *blk_m, uint64_t *blk_g,
                              uint32_t *status, int lane) {
    const uint32_t n = sum_cnt[top];
    bool ok = n == static_cast<uint32_t>(kk);
    if (ok) {
        const uint64_t q = lane < kk ? sum_q[top * kWave + lane] : 0ull;
        const uint32_t rk = lane < kk ? sum_rank[top * kWave + lane] : 0u;
        // column j of the rank-ordered word matrix (words have no bit at or above kk, rows at or above
        // kk are zero)
        uint64_t col = transpose64(q, lane);
        uint64_t m = lane < kk ? (1ull << lane) : 0ull;
        // Branch-free steps (columns at or above kk are zero, so they never take an update): the
        // chain from one step to the next is a lane read, the column's lowest bit and a masked XOR.
        // Step b clears its winner's bit from every later column, so column b holds no earlier
        // winner's bit by its own step: its lowest bit is the lowest untaken one, and column b at
        // the end (no later step changes it) gives phase b's winner.  A column without a bit at its
        // step (zero at the end) means the block fails.
        for (int b = 0; b < kk; ++b) {
            const uint64_t cb = rl64(col, b);
            const uint64_t rbit = lowbit(cb);
            const uint64_t mb = rl64(m, b);
            const bool hit = lane > b && (col & rbit) != 0;
            m ^= hit ? mb : 0ull;
            col ^= hit ? cb : 0ull;
        }
        ok = !__ballot(lane < kk && col == 0);
        const uint64_t winbit = lane < kk ? lowbit(col) : 0ull;  // lane i: phase i's winner as a bit of the rank positions
        const int win = winbit ? __ffsll(static_cast<unsigned long long>(winbit)) - 1 : 0;
        if (ok) {
            const uint32_t qlo = static_cast<uint32_t>(__shfl(static_cast<int>(static_cast<uint32_t>(q)), win));
            const uint32_t qhi = static_cast<uint32_t>(__shfl(static_cast<int>(static_cast<uint32_t>(q >> 32)), win));
            const uint64_t qi = (static_cast<uint64_t>(qhi) << 32) | qlo;
            const uint32_t ri = static_cast<uint32_t>(__shfl(static_cast<int>(rk), win));
            // lane i: bit l = <C_{e+i}, final S_{e+l}>, bit win_i of column l: row win_i of the
            // transposed columns (columns at or above kk are zero)
            const uint64_t rows = transpose64(col, lane);
            const uint32_t plo = static_cast<uint32_t>(__shfl(static_cast<int>(static_cast<uint32_t>(rows)), win));
            const uint32_t phi = static_cast<uint32_t>(__shfl(static_cast<int>(static_cast<uint32_t>(rows >> 32)), win));
            const uint64_t prod = lane < kk ? (static_cast<uint64_t>(phi) << 32) | plo : 0ull;
            const bool odd = lane >= kk || ((prod >> lane) & 1ull);
            if (__ballot(!odd)) ok = false;
            if (lane < kk) {
                sel[lane] = ri;
                blk_q[lane] = qi;
                blk_m[lane] = m;
                blk_g[lane] = prod & ((1ull << lane) - 1);
            }
        }
    }
    if (!ok) {
        if (lane == 0) atomicOr(status, 1u);
        if (lane < kk) sel[lane] = kNoRank;
    }
}

__device__ __forceinline__ void block_finalize(const BlockArgs &a, int lane) {
    finalize_core(a.kk, a.sum_cnt, a.sum_rank, a.sum_q, a.lvl_slot[a.levels - 1], a.sel + a.e, a.blk_q, a.blk_m,
                  a.blk_g, a.status, lane);
}

// Position of the k-th (from 0) set bit of w, which has more than k set bits: halving by popcounts.
__device__ __forceinline__ int select_bit(uint64_t w, uint32_t k) {
    int pos = 0;
#pragma unroll
    for (int s = 32; s >= 2; s >>= 1) {
        const uint32_t c = static_cast<uint32_t>(__popcll(w & ((1ull << s) - 1)));
        if (k >= c) {
            k -= c;
            w >>= s;
            pos += s;
        }
    }
    return pos + (k >= static_cast<uint32_t>(w & 1ull) ? 1 : 0);
}

// First level of the block's greedy basis for the ranks [r0, r0 + 64 * nwords) (r0 a multiple of
// 64, nwords <= 64 * kR) into one slot; only the words the tree pass emitted take part (the others
// cannot be in the basis), and the bitmap words read are cleared for the next block.  The bitmap is
// read in rounds of 64 words, all rounds' loads in flight together; prefix sums of the popcounts
// list the emitted ranks in rank order (s_words / s_offs: 64 * kR entries of LDS), and they are
// inserted 64 at a time, each batch's words loaded while the one before it is inserted.
template <int kR>
__device__ __forceinline__ void basis_chunk(uint64_t *qbits, const uint64_t *Qr, int64_t r0, int nwords, int64_t slot,
                                            int kk, uint32_t *sum_cnt, uint32_t *sum_rank, uint64_t *sum_q,
                                            uint64_t *s_words, uint32_t *s_offs, int lane) {
    const int64_t wbase = r0 >> 6;
    const int rounds = (nwords + kWave - 1) / kWave;
    uint64_t word[kR];
#pragma unroll
    for (int j = 0; j < kR; ++j) {
        const int w = j * kWave + lane;
        word[j] = (j < rounds && w < nwords) ? qbits[wbase + w] : 0ull;
    }
    uint32_t total = 0;
#pragma unroll
    for (int j = 0; j < kR; ++j) {
        if (j >= rounds) break;
        if (word[j]) qbits[wbase + j * kWave + lane] = 0;  // ready for the next block
        const uint32_t own = static_cast<uint32_t>(__popcll(word[j]));
        uint32_t incl = own;
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            const uint32_t t = __shfl_up(incl, o);
            if (lane >= o) incl += t;
        }
        s_words[j * kWave + lane] = word[j];
        s_offs[j * kWave + lane] = total + incl - own;
        total += static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(incl), kWave - 1));
    }
    __builtin_amdgcn_wave_barrier();
    const int nx = rounds * kWave;
    // entry t: the last word whose first entry is at or before t owns it
    auto fetch = [&](uint32_t t, uint64_t &q, uint32_t &rank) {
        q = 0;
        rank = 0;
        if (t < total) {
            int lo = 0, hi = nx - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (s_offs[mid] <= t) lo = mid;
                else hi = mid - 1;
            }
            rank = static_cast<uint32_t>(r0 + static_cast<int64_t>(lo) * kWave + select_bit(s_words[lo], t - s_offs[lo]));
            q = Qr[rank];
        }
    };
    uint64_t qn;
    uint32_t rn;
    fetch(static_cast<uint32_t>(lane), qn, rn);
    uint64_t piv = 0, P = 0;
    uint32_t cnt = 0;
    for (uint32_t b0 = 0; b0 < total && cnt < static_cast<uint32_t>(kk); b0 += kWave) {
        const uint64_t q = qn;
        const uint32_t rank = rn;
        fetch(b0 + kWave + static_cast<uint32_t>(lane), qn, rn);
        basis_insert(piv, P, cnt, q, rank, sum_rank + slot * kWave, sum_q + slot * kWave, lane);
    }
    if (lane == 0) sum_cnt[slot] = cnt;
}

// Merges the `members` bases (at most 64) in slots s0 .. s0+members-1 into slot `slot`.  Their
// entries, concatenated in order, are in rank order.  All of them are first copied into LDS (one
// coalesced load per base, sixteen bases in flight: the bases were written by the previous launch,
// often through another XCD's L2, so a load per batch would put one memory latency on every batch of
// the serial insertion), then inserted 64 at a time.  The kMergeWaves waves of the workgroup share the
// loads; wave 0 inserts (the others return false).  tab: kMergeTab words of reduction tables, st_q /
// st_r: kMergeEntries entries.
constexpr int kMergeWaves = 4;  // waves per merge workgroup (one merge group, 50 KB of LDS)
constexpr int kMergeEntries = 64 * kWave;
__device__ __forceinline__ bool merge_group(int64_t s0, int members, int64_t slot, int kk, uint32_t *sum_cnt,
                                            uint32_t *sum_rank, uint64_t *sum_q, uint64_t *tab, uint64_t *st_q,
                                            uint32_t *st_r, int lane, int wv) {
    const uint32_t own = lane < members ? sum_cnt[s0 + lane] : 0u;
    uint32_t incl = own;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const uint32_t t = __shfl_up(incl, o);
        if (lane >= o) incl += t;
    }
    const uint32_t total = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(incl), kWave - 1));
    const uint32_t excl = incl - own;
    constexpr int kInFlight = 16;
    for (int m0 = wv * kInFlight; m0 < members; m0 += kMergeWaves * kInFlight) {
        uint64_t q[kInFlight];
        uint32_t r[kInFlight];
#pragma unroll
        for (int u = 0; u < kInFlight; ++u) {
            const int m = m0 + u;
            const uint32_t c = m < members ? static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(own), m)) : 0u;
            q[u] = 0;
            r[u] = 0;
            if (static_cast<uint32_t>(lane) < c) {
                const int64_t at = (s0 + m) * kWave + lane;
                q[u] = sum_q[at];
                r[u] = sum_rank[at];
            }
        }
#pragma unroll
        for (int u = 0; u < kInFlight; ++u) {
            const int m = m0 + u;
            const uint32_t c = m < members ? static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(own), m)) : 0u;
            if (static_cast<uint32_t>(lane) < c) {
                const uint32_t o = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(excl), m)) + lane;
                st_q[o] = q[u];
                st_r[o] = r[u];
            }
        }
    }
    __syncthreads();  // every wave's share is in LDS; wave 0 inserts
    if (wv != 0) return false;
    uint64_t piv = 0, P = 0;
    uint32_t cnt = 0;
    bool stale = true;
    for (uint32_t b0 = 0; b0 < total && cnt < static_cast<uint32_t>(kk); b0 += kWave) {
        const uint32_t t = b0 + lane;
        const uint64_t q = t < total ? st_q[t] : 0ull;
        const uint32_t r = t < total ? st_r[t] : 0u;
        basis_insert_tab(piv, P, cnt, stale, tab, q, r, sum_rank + slot * kWave, sum_q + slot * kWave, lane);
    }
    if (lane == 0) sum_cnt[slot] = cnt;
    return true;
}

// Merges the `members` bases (at most 64) in slots s0 .. s0+members-1 into slot `slot` with one
// wave.  Their entries, concatenated in order, are in rank order; they are inserted 64 at a time,
// each batch loaded while the one before it is inserted.  A merge reaches kk pivots after a few
// batches, so only those are read.  offs: 64 words of LDS (the members' first entries), tab:
// kMergeTab words of reduction tables.
__device__ __forceinline__ void merge_stream(int64_t s0, int members, int64_t slot, int kk, uint32_t *sum_cnt,
                                             uint32_t *sum_rank, uint64_t *sum_q, uint64_t *tab, uint32_t *offs,
                                             int lane) {
    const uint32_t own = lane < members ? sum_cnt[s0 + lane] : 0u;
    uint32_t incl = own;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const uint32_t t = __shfl_up(incl, o);
        if (lane >= o) incl += t;
    }
    const uint32_t total = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(incl), kWave - 1));
    offs[lane] = incl - own;  // the lanes past the members hold total
    __builtin_amdgcn_wave_barrier();
    // entry t: the last member whose first entry is at or before t owns it
    auto fetch = [&](uint32_t t, uint64_t &q, uint32_t &r) {
        q = 0;
        r = 0;
        if (t < total) {
            int lo = 0, hi = kWave - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (offs[mid] <= t) lo = mid;
                else hi = mid - 1;
            }
            const int64_t at = (s0 + lo) * kWave + (t - offs[lo]);
            q = sum_q[at];
            r = sum_rank[at];
        }
    };
    uint64_t qn;
    uint32_t rn;
    fetch(static_cast<uint32_t>(lane), qn, rn);
    uint64_t piv = 0, P = 0;
    uint32_t cnt = 0;
    bool stale = true;
    for (uint32_t b0 = 0; b0 < total && cnt < static_cast<uint32_t>(kk); b0 += kWave) {
        const uint64_t q = qn;
        const uint32_t r = rn;
        fetch(b0 + kWave + lane, qn, rn);
        basis_insert_tab(piv, P, cnt, stale, tab, q, r, sum_rank + slot * kWave, sum_q + slot * kWave, lane);
    }
    if (lane == 0) sum_cnt[slot] = cnt;
}

// Wave c: the bitmap words [c * chunk_words, (c + 1) * chunk_words) (chunk_words a multiple of 64, at
// most 64 * kMaxRounds) -> slot c of the first level; with one level the block's finalisation, and with
// one merge above it that merge and the finalisation (levels above that: k_block_cascade).
__global__ __launch_bounds__(256) void k_block_basis(BlockArgs a) {
    __shared__ uint64_t s_words[4][kMaxRounds * kWave];
    __shared__ uint32_t s_offs[4][kMaxRounds * kWave];
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = wave_id();
    const int64_t c = static_cast<int64_t>(blockIdx.x) * 4 + wv;
    if (c >= a.lvl_n[0]) return;
    const int64_t w0 = c * a.chunk_words;
    const int64_t nwords = (a.K + kWave - 1) / kWave - w0;
    basis_chunk<kMaxRounds>(a.qbits, a.Qr, w0 * kWave, static_cast<int>(nwords < a.chunk_words ? nwords : a.chunk_words),
                            a.lvl_slot[0] + c, a.kk, a.sum_cnt, a.sum_rank, a.sum_q, s_words[wv], s_offs[wv], lane);
    if (a.levels == 1) {
        __threadfence();
        block_finalize(a, lane);
    } else if (a.levels == 2) {
        // One merge above the first level (a few dozen bases): the last wave to finish merges them and
        // finalises the block, so the block takes no launch of its own for the merge.  Each wave
        // publishes its base with an agent-scope release fence before it counts itself; the last
        // arrival clears the counter for the next block and takes an acquire fence before it reads the
        // others' bases.
        const int64_t dst = a.lvl_slot[1];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        uint32_t before = 0;
        if (lane == 0) before = atomicAdd(a.arrive + dst, 1u);
        before = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(before), 0));
        if (before + 1 != static_cast<uint32_t>(a.lvl_n[0])) return;
        if (lane == 0) atomicExch(a.arrive + dst, 0u);
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        // this wave's bitmap staging is free: its first 256 words hold the reduction tables
        merge_stream(a.lvl_slot[0], static_cast<int>(a.lvl_n[0]), dst, a.kk, a.sum_cnt, a.sum_rank, a.sum_q,
                     s_words[wv], s_offs[wv], lane);
        __threadfence();
        block_finalize(a, lane);
    }
}

// Every merge level of the block's greedy basis in one launch of one-wave workgroups.  Workgroup g
// merges first-level group g (written by the previous launch); from there on, the last member of
// each merge group to finish merges that group: each wave publishes its merged base with an
// agent-scope release fence before it counts itself in the group's arrival counter, and the last
// arrival clears the counter for the next block and takes an acquire fence before it reads the
// others' bases.  The wave that completes the top level finalises the block.  (Used when the block
// has more than one merge level; the first level stays a launch of its own then, since fences in its
// hundreds of waves, each writing back and invalidating its L2, cost more than the launch.)
__global__ __launch_bounds__(64) void k_block_cascade(BlockArgs a) {
    __shared__ uint32_t s_offs[kWave];
    __shared__ uint64_t s_tab[kMergeTab];
    const int lane = threadIdx.x & (kWave - 1);
    int64_t c = blockIdx.x;
    if (c >= a.lvl_n[1]) return;
    for (int L = 0; L + 1 < a.levels; ++L) {
        const int64_t g = L == 0 ? c : c / a.group;
        const int64_t rest = a.lvl_n[L] - g * a.group;
        const int members = static_cast<int>(rest < a.group ? rest : a.group);
        const int64_t dst = a.lvl_slot[L + 1] + g;
        if (L > 0) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            uint32_t before = 0;
            if (lane == 0) before = atomicAdd(a.arrive + dst, 1u);
            before = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(before), 0));
            if (before + 1 != static_cast<uint32_t>(members)) return;
            if (lane == 0) atomicExch(a.arrive + dst, 0u);
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
        merge_stream(a.lvl_slot[L] + g * a.group, members, dst, a.kk, a.sum_cnt, a.sum_rank, a.sum_q, s_tab,
                     s_offs, lane);
        c = g;
    }
    __threadfence();
    block_finalize(a, lane);
}

// C of one chosen candidate into LDS (W words, cleared by the caller): the coordinates of the edges
// on root -> u and root -> v, found by testing the positions of the root's tree against the DFS
// intervals of u and v, and the candidate's own edge.  A parent precedes its child (checked at
// create), so only the positions up to max(u, v) are tested.  Coordinates are counted from cb.
__device__ void cycle_into_lds(const int64_t *tree_off, const uint32_t *pos_pre, const uint32_t *pos_end,
                               const int32_t *pos_coord, uint32_t root, uint32_t cu, uint32_t cv, int32_t ccoord,
                               int64_t cb, uint64_t *Cl) {
    constexpr int kB = 8;
    const int tid = threadIdx.x;
    const int64_t off = tree_off[root];
    const uint32_t npos = static_cast<uint32_t>(tree_off[root + 1] - off);
    const uint32_t pu = pos_pre[off + cu];
    const uint32_t pv = pos_pre[off + cv];
    const uint32_t last = cu > cv ? cu : cv;
    const uint32_t lim = last + 1 < npos ? last + 1 : npos;
    if (tid == 0 && ccoord >= 0) {
        const int64_t c = ccoord - cb;
        atomicOr(reinterpret_cast<unsigned long long *>(&Cl[c >> 6]), 1ull << (c & 63));
    }
    for (uint32_t x0 = 1 + tid; x0 < lim; x0 += kB * blockDim.x) {
        uint32_t lo[kB], hi[kB];
        int32_t c[kB];
#pragma unroll
        for (int k = 0; k < kB; ++k) {
            const uint32_t x = x0 + k * blockDim.x;
            const uint32_t at = x < lim ? x : 0u;
            lo[k] = pos_pre[off + at];
            hi[k] = pos_end[off + at];
            c[k] = x < lim ? pos_coord[off + at] : -1;
        }
#pragma unroll
        for (int k = 0; k < kB; ++k) {
            const bool on = (lo[k] <= pu && pu < hi[k]) || (lo[k] <= pv && pv < hi[k]);
            if (on && c[k] >= 0) {
                const int64_t cl = c[k] - cb;
                atomicOr(reinterpret_cast<unsigned long long *>(&Cl[cl >> 6]), 1ull << (cl & 63));
            }
        }
    }
}

// C_{e+i} of phase e+i's winner and its nonzero words below e.  One workgroup; Cl: W words of LDS,
// cnt: one LDS counter.
template <typename A>
__device__ __forceinline__ void cycles_body(const A &a, const BlkView &v, int i, uint64_t *Cl, uint32_t *cnt) {
    const int tid = threadIdx.x;
    for (int w = tid; w < v.W; w += blockDim.x) Cl[w] = 0;
    if (tid == 0) *cnt = 0;
    __syncthreads();
    const uint32_t rank = v.sel[v.e + i];
    if (rank < a.K)
        cycle_into_lds(a.tree_off, a.pos_pre, a.pos_end, a.pos_coord, a.cand_root[rank], a.cand_pu[rank],
                       a.cand_pv[rank], a.cand_coord[rank], v.cb, Cl);
    __syncthreads();
    const int32_t dlim = (v.e + kWave - 1) / kWave;
    uint64_t *Cout = v.C_blk + static_cast<int64_t>(i) * v.W;
    uint32_t *Dout = v.D_blk + static_cast<int64_t>(i) * v.W;
    for (int w = tid; w < v.W; w += blockDim.x) {
        const uint64_t x = Cl[w];
        Cout[w] = x;
        if (v.C_store) v.C_store[(static_cast<int64_t>(v.e) + i) * v.W + w] = x;
        if (w < dlim && x) Dout[atomicAdd(cnt, 1u)] = static_cast<uint32_t>(w);
    }
    __syncthreads();
    if (tid == 0) v.Dcnt_blk[i] = *cnt;
}

// The block's final support vectors S_{e+i} = sum over l of m_i[l] T_l on word w (a 64 x 64 product
// over GF(2)): lane l loads T_l's word w (with its own bit), lane i accumulates the rows m_i selects.
// Sf_blk[w][i] holds the result.  One wave per word.
__device__ __forceinline__ void final_body(const BlkView &v, int64_t w, int lane) {
    const int64_t j = static_cast<int64_t>(v.e) + lane;
    uint64_t x = 0;
    if (lane < v.kk) {
        x = v.Sw[w * v.npad + j];
        if ((j >> 6) == w) x |= 1ull << (j & 63);
    }
    const uint64_t m = lane < v.kk ? v.blk_m[lane] : 0ull;
    uint64_t acc = 0;
    for (int l = 0; l < v.kk; ++l) {
        const uint64_t t = rl64(x, l);
        if ((m >> l) & 1ull) acc ^= t;
    }
    v.Sf_blk[w * kBlockK + lane] = acc;
}

// After the block's finalisation, in one launch: workgroup i < kk builds phase e+i's cycle vector
// (cycles_body), the others the final support vectors, one word per wave (final_body).
__global__ __launch_bounds__(256) void k_block_cycles_final(BlockArgs a) {
    extern __shared__ uint64_t lds[];
    __shared__ uint32_t cnt;
    if (static_cast<int>(blockIdx.x) < a.kk) {
        cycles_body(a, view_of(a), blockIdx.x, lds, &cnt);
        return;
    }
    const int64_t w = static_cast<int64_t>(blockIdx.x - a.kk) * 4 + wave_id();
    const int64_t wl = (static_cast<int64_t>(a.e) + a.kk + kWave - 1) / kWave;
    if (w >= wl) return;
    final_body(view_of(a), w, threadIdx.x & (kWave - 1));
}

// Ct[64w + b] bit i = bit b of word w of C_{e+i}: one wave per word.
__device__ __forceinline__ void pairs_transpose(const BlkView &v, int64_t w, int lane) {
    const uint64_t x = lane < v.kk ? v.C_blk[static_cast<int64_t>(lane) * v.W + w] : 0ull;
    v.Ct[w * kWave + lane] = transpose64(x, lane);
}

// The words below e where some C_{e+i} is nonzero, in ascending order, each with its (i, word of
// C_{e+i}) pairs and the prefix sums of the pair counts, so that the pairs of consecutive listed
// words are contiguous.  One workgroup; lds: ceil(e/64) words + 2 * blockDim.x u32.
__device__ __forceinline__ void pairs_list(const BlkView &v, uint64_t *lds, uint32_t *nlist, uint32_t *npairs) {
    __shared__ uint32_t pre[kBlockK + 1];  // prefix sums of the nonzero-word counts of C_{e+i}
    __shared__ uint32_t wsum[2][8];        // per-wave totals of the scan below
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1), wid = wave_id();
    // U[w]: bit i set when C_{e+i} is nonzero at word w (w below e)
    const int32_t dlim = (v.e + kWave - 1) / kWave;
    uint64_t *U = lds;
    uint32_t *part = reinterpret_cast<uint32_t *>(U + dlim);  // [2 * blockDim.x]: words, pairs per thread
    for (int x = tid; x < dlim; x += blockDim.x) U[x] = 0;
    if (wid == 0) {
        const uint32_t c = lane < v.kk ? v.Dcnt_blk[lane] : 0u;
        uint32_t incl = c;
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            const uint32_t t = __shfl_up(incl, o);
            if (lane >= o) incl += t;
        }
        pre[lane + 1] = incl;
        if (lane == 0) pre[0] = 0;
    }
    __syncthreads();
    // every (i, listed word of C_{e+i}) entry at once: the loads are independent
    const uint32_t total = pre[v.kk];
    for (uint32_t idx = tid; idx < total; idx += blockDim.x) {
        int lo = 0, hi = v.kk - 1;  // the last i whose entries start at or before idx
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (pre[mid] <= idx) lo = mid;
            else hi = mid - 1;
        }
        const uint32_t w = v.D_blk[static_cast<int64_t>(lo) * v.W + (idx - pre[lo])];
        atomicOr(reinterpret_cast<unsigned long long *>(&U[w]), 1ull << lo);
    }
    __syncthreads();
    const int per = (dlim + static_cast<int>(blockDim.x) - 1) / static_cast<int>(blockDim.x);
    const int w0 = tid * per, w1 = w0 + per < dlim ? w0 + per : dlim;
    uint32_t nw = 0, np = 0;
    for (int w = w0; w < w1; ++w)
        if (U[w]) {
            ++nw;
            np += static_cast<uint32_t>(__popcll(U[w]));
        }
    // exclusive scans of (words, pairs) over the threads: within each wave, then across the waves
    uint32_t iw = nw, ip = np;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const uint32_t tw = __shfl_up(iw, o), tp = __shfl_up(ip, o);
        if (lane >= o) {
            iw += tw;
            ip += tp;
        }
    }
    if (lane == kWave - 1) {
        wsum[0][wid] = iw;
        wsum[1][wid] = ip;
    }
    __syncthreads();
    uint32_t bw = 0, bp = 0, tw = 0, tp = 0;
    const int nwav = static_cast<int>(blockDim.x) / kWave;
    for (int k = 0; k < nwav; ++k) {
        if (k < wid) {
            bw += wsum[0][k];
            bp += wsum[1][k];
        }
        tw += wsum[0][k];
        tp += wsum[1][k];
    }
    part[tid] = bw + iw - nw;
    part[blockDim.x + tid] = bp + ip - np;
    if (tid == 0) {
        *nlist = tw;
        *npairs = tp;
    }
    __syncthreads();
    uint32_t x = part[tid], off = part[blockDim.x + tid];
    for (int w = w0; w < w1; ++w) {
        uint64_t m = U[w];
        if (!m) continue;
        v.pw_list[x] = static_cast<uint32_t>(w);
        v.pw_off[x] = off;
        ++x;
        while (m) {
            const int i = __ffsll(static_cast<unsigned long long>(m)) - 1;
            m &= m - 1;
            v.pair_i[off] = static_cast<uint32_t>(i);
            v.pair_c[off] = static_cast<uint64_t>(w);  // the word; its value of C_{e+i} follows below
            ++off;
        }
    }
    if (tid == 0) {
        v.pw_off[*nlist] = *npairs;
        *v.pw_cnt = *nlist;
    }
    __syncthreads();  // the pairs above are visible to the whole workgroup
    // every pair's word of C at once, four independent loads in flight per thread
    constexpr int kP = 4;
    for (uint32_t q0 = tid; q0 < tp; q0 += kP * blockDim.x) {
        uint32_t pi[kP];
        uint64_t pw[kP], pc[kP];
#pragma unroll
        for (int k = 0; k < kP; ++k) {
            const uint32_t q = q0 + k * blockDim.x;
            pi[k] = q < tp ? v.pair_i[q] : 0u;
            pw[k] = q < tp ? v.pair_c[q] : 0ull;
        }
#pragma unroll
        for (int k = 0; k < kP; ++k) {
            const int64_t at = static_cast<int64_t>(pi[k]) * v.W + static_cast<int64_t>(pw[k]);
            pc[k] = q0 + k * blockDim.x < tp ? v.C_blk[at] : 0ull;
        }
#pragma unroll
        for (int k = 0; k < kP; ++k)
            if (q0 + k * blockDim.x < tp) v.pair_c[q0 + k * blockDim.x] = pc[k];
    }
}

// Inputs of the update that every column shares: workgroup 0 lists the words, the others
// transpose the block's cycle vectors (4 words each).
__global__ __launch_bounds__(256) void k_block_pairs(BlockArgs a) {
    extern __shared__ uint64_t lds[];
    __shared__ uint32_t nlist, npairs;
    const BlkView v = view_of(a);
    if (blockIdx.x > 0) {
        const int64_t w = static_cast<int64_t>(blockIdx.x - 1) * 4 + wave_id();
        if (w < v.W) pairs_transpose(v, w, threadIdx.x & (kWave - 1));
        return;
    }
    pairs_list(v, lds, &nlist, &npairs);
}

// The block's update of a later column j >= e+kk is S_j ^= sum_i b_i S_{e+i} with
// b_i = <C_{e+i}, S_j> ^ sum_{l<i} b_l <C_{e+i}, S_{e+l}>.  S_j has support below e besides its own
// bit, so the products read only the listed words (where some C_{e+i} is nonzero below e) plus bit j
// of each C (Ct[j]).  Column tile t: lane l of every wave takes column j = e + kk + 64 t + l, the
// kDotWaves waves take runs of kB listed words in turn (a shorter chain of dependent loads per lane),
// and wave 0 combines their partial products and computes b into Bc[j].  The list and its pairs are
// staged in LDS in segments of at most kSegWords words and kSegPairs pairs.
// LDS: G rows [64] | pair words [kSegPairs] | pair phases [kSegPairs] | listed words [kSegWords] |
// their pair offsets [kSegWords + 1].
constexpr int kSegWords = 1024;
constexpr int kSegPairs = 2048;
constexpr int kDotWaves = 4;      // waves per product workgroup
constexpr int kDotCols = kWave;   // later columns per product tile

// Column j's products d (bit i: <C_{e+i}, S_j>) into its coefficients b (the recurrence over the block's
// phases with the rows G), written to Bc[j]; the columns with b != 0 go to the update list (one atomic
// per wave).  Returns the bytes the wave's columns read (in lane 0).  One wave, lane = column.
__device__ __forceinline__ uint64_t finish_column(const BlkView &v, const uint64_t *G, int64_t j, bool col,
                                                  uint64_t d, uint32_t nl, int lane) {
    uint64_t nbytes = 0, bcol = 0;
    if (col) {
        uint64_t b = 0;
        for (int i = 0; i < v.kk; ++i) {
            const uint64_t bi = ((d >> i) ^ static_cast<uint64_t>(__popcll(b & G[i]))) & 1ull;
            b |= bi << i;
        }
        v.Bc[j] = b;
        bcol = b;
        nbytes = 2 + 8ull * nl;
    }
    if (v.ulist) {  // list the columns that take an update, one atomic per wave
        const bool upd = col && bcol != 0;
        const uint64_t m = __ballot(upd);
        if (m) {
            uint32_t base = 0;
            if (lane == 0) base = atomicAdd(v.ucount, static_cast<uint32_t>(__popcll(m)));
            base = static_cast<uint32_t>(__shfl(static_cast<int>(base), 0));
            if (upd) v.ulist[base + __popcll(m & ((1ull << lane) - 1))] = static_cast<uint32_t>(j);
        }
    }
    for (int o = kWave / 2; o > 0; o >>= 1) nbytes += __shfl_down(nbytes, o);
    return nbytes;  // the wave's total in lane 0
}

// The later columns when the block's word list is short (at most kShortList words, their pairs in one
// LDS segment), as on sparse graphs: the list, its pairs and the rows G are staged once per workgroup,
// then each wave takes whole 64-column tiles in turn (lane = column: one load per listed word, all in
// flight together), so no workgroup repeats the staging per tile.  Returns false, having done nothing,
// for a longer list.  LDS as dot_body's.
constexpr uint32_t kShortList = 48;
__device__ __forceinline__ bool dot_short(const BlkView &v, uint64_t *lds, unsigned long long *upd_bytes) {
    constexpr int kB = 16;
    const uint32_t nl = uniform32(*v.pw_cnt);
    if (nl > kShortList) return false;
    const uint32_t np = uniform32(v.pw_off[nl]) - uniform32(v.pw_off[0]);
    if (np > static_cast<uint32_t>(kSegPairs)) return false;
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wv = tid / kWave;
    const int nwv = static_cast<int>(blockDim.x) / kWave;
    uint64_t *G = lds;
    uint64_t *Pc = G + kBlockK;
    uint32_t *Pi = reinterpret_cast<uint32_t *>(Pc + kSegPairs);
    uint32_t *Lw = Pi + kSegPairs;
    uint32_t *Lo = Lw + kSegWords;
    const uint32_t p0 = nl ? v.pw_off[0] : 0u;
    for (int x = tid; x < kBlockK; x += blockDim.x) G[x] = x < v.kk ? v.blk_g[x] : 0ull;
    for (uint32_t x = tid; x < nl; x += blockDim.x) {
        Lw[x] = v.pw_list[x];
        Lo[x] = v.pw_off[x] - p0;
    }
    if (tid == 0) Lo[nl] = np;
    for (uint32_t q = tid; q < np; q += blockDim.x) {
        Pc[q] = v.pair_c[p0 + q];
        Pi[q] = v.pair_i[p0 + q];
    }
    __syncthreads();
    const int64_t later = static_cast<int64_t>(v.N) - v.e - v.kk;
    const int64_t tiles = (later + kDotCols - 1) / kDotCols;
    uint64_t nbytes = 0;
    for (int64_t t = static_cast<int64_t>(blockIdx.x) * nwv + wv; t < tiles;
         t += static_cast<int64_t>(gridDim.x) * nwv) {
        const int64_t j = static_cast<int64_t>(v.e) + v.kk + t * kDotCols + lane;
        const bool col = j < v.N;
        uint64_t d = col ? v.Ct[j] : 0ull;
        for (uint32_t x0 = 0; x0 < nl; x0 += kB) {
            uint64_t sv[kB];
#pragma unroll
            for (int k = 0; k < kB; ++k) {
                const uint32_t x = x0 + k < nl ? x0 + k : x0;
                sv[k] = col ? v.Sw[static_cast<int64_t>(Lw[x]) * v.npad + j] : 0ull;
            }
#pragma unroll
            for (int k = 0; k < kB; ++k) {
                if (x0 + k >= nl) break;
                for (uint32_t q = Lo[x0 + k]; q < Lo[x0 + k + 1]; ++q)
                    d ^= static_cast<uint64_t>(__popcll(Pc[q] & sv[k]) & 1) << Pi[q];
            }
        }
        nbytes += finish_column(v, G, j, col, d, nl, lane);
    }
    if (lane == 0 && nbytes) add_bytes(upd_bytes, nbytes);
    return true;
}
__device__ __forceinline__ void dot_body(const BlkView &v, int64_t tile, uint64_t *lds, uint32_t *seg_end,
                                         unsigned long long *upd_bytes) {
    constexpr int kB = 16;
    static_assert(kDotWaves * kWave <= kSegPairs, "partial products fit the pair words");
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wv = tid / kWave;
    const int kk = v.kk;
    uint64_t *G = lds;
    uint64_t *Pc = G + kBlockK;
    uint32_t *Pi = reinterpret_cast<uint32_t *>(Pc + kSegPairs);
    uint32_t *Lw = Pi + kSegPairs;
    uint32_t *Lo = Lw + kSegWords;
    for (int x = tid; x < kBlockK; x += blockDim.x) G[x] = x < kk ? v.blk_g[x] : 0ull;
    const int64_t j = static_cast<int64_t>(v.e) + kk + tile * kDotCols + lane;
    const bool col = j < v.N;
    uint64_t d = col && wv == 0 ? v.Ct[j] : 0ull;
    const uint32_t nl = *v.pw_cnt;
    for (uint32_t s0 = 0; s0 < nl;) {
        __syncthreads();
        // the pair offsets of the next kSegWords listed words, staged once, then the longest run of
        // them from s0 whose pairs fit (searched in LDS)
        const uint32_t hw = s0 + kSegWords < nl ? s0 + kSegWords : nl;
        for (uint32_t x = tid; x <= hw - s0; x += blockDim.x) Lo[x] = v.pw_off[s0 + x];
        __syncthreads();
        if (tid == 0) {
            const uint32_t q0 = Lo[0];
            uint32_t lo = s0 + 1, hi = hw;
            while (lo < hi) {
                const uint32_t mid = (lo + hi + 1) / 2;
                if (Lo[mid - s0] - q0 <= static_cast<uint32_t>(kSegPairs)) lo = mid;
                else hi = mid - 1;
            }
            *seg_end = lo;
        }
        __syncthreads();
        const uint32_t s1 = *seg_end;
        const uint32_t p0 = Lo[0], p1 = Lo[s1 - s0];
        __syncthreads();  // every thread holds p0 and p1 before the offsets are rebased
        for (uint32_t x = tid; x < s1 - s0; x += blockDim.x) {
            Lw[x] = v.pw_list[s0 + x];
            Lo[x] -= p0;
        }
        if (tid == 0) Lo[s1 - s0] = p1 - p0;
        for (uint32_t q = tid; q < p1 - p0; q += blockDim.x) {
            Pc[q] = v.pair_c[p0 + q];
            Pi[q] = v.pair_i[p0 + q];
        }
        __syncthreads();
        if (col) {
            for (uint32_t x0 = wv * kB; x0 < s1 - s0; x0 += kDotWaves * kB) {
                uint64_t sv[kB];
#pragma unroll
                for (int k = 0; k < kB; ++k) {
                    const uint32_t x = x0 + k < s1 - s0 ? x0 + k : x0;
                    sv[k] = v.Sw[static_cast<int64_t>(Lw[x]) * v.npad + j];
                }
#pragma unroll
                for (int k = 0; k < kB; ++k) {
                    if (x0 + k >= s1 - s0) break;
                    for (uint32_t q = Lo[x0 + k]; q < Lo[x0 + k + 1]; ++q)
                        d ^= static_cast<uint64_t>(__popcll(Pc[q] & sv[k]) & 1) << Pi[q];
                }
            }
        }
        s0 = s1;
    }
    __syncthreads();  // the pair words are no longer read: they take the waves' partial products
    Pc[tid] = d;
    __syncthreads();
    if (wv == 0) {
        for (int w = 1; w < kDotWaves; ++w) d ^= Pc[w * kWave + lane];
        const uint64_t nbytes = finish_column(v, G, j, col, d, nl, lane);
        if (lane == 0 && nbytes) add_bytes(upd_bytes, nbytes);
    }
    __syncthreads();  // the LDS is free for the next tile
}

__global__ __launch_bounds__(kDotWaves * kWave) void k_block_dot(BlockArgs a) {
    extern __shared__ uint64_t lds[];
    __shared__ uint32_t seg_end;
    const BlkView v = view_of(a);
    if (dot_short(v, lds, a.upd_bytes)) return;
    const int64_t tiles = (static_cast<int64_t>(v.N) - v.e - v.kk + kDotCols - 1) / kDotCols;
    for (int64_t t = blockIdx.x; t < tiles; t += gridDim.x) dot_body(v, t, lds, &seg_end, a.upd_bytes);
}

// The XOR update of tile (x, y) of kTileDirect words, combining the block's final vectors staged in LDS
// (T[t][i]) directly: 64 masked XORs per word.  Cheaper than the tables of xor_body when a step's tiles
// hold few columns each (the batched form's small components).  Column tile x covers the listed later
// columns ulist[256 x ..] (those with nonzero coefficients); tile 0 also writes the block's own columns.
__device__ __forceinline__ void xor_direct(const BlkView &v, int64_t xt, int32_t yt, uint64_t *lds,
                                           unsigned long long *upd_bytes) {
    constexpr int kW = 8;
    const int tid = threadIdx.x;
    const int kk = v.kk;
    const int32_t wl = (v.e + kk + kWave - 1) / kWave;
    const int32_t w0 = yt * kTileDirect;
    const int tn = wl - w0 < kTileDirect ? wl - w0 : kTileDirect;
    const int64_t J = uniform32(*v.ucount);
    if (xt > 0 && xt * blockDim.x >= J) return;  // no listed column in this tile
    uint64_t *T = lds;  // [kTileDirect][64]
    for (int x = tid; x < kBlockK * kTileDirect; x += blockDim.x) {
        const int t = x / kBlockK, i = x % kBlockK;
        T[x] = (i < kk && t < tn) ? v.Sf_blk[static_cast<int64_t>(w0 + t) * kBlockK + i] : 0ull;
    }
    __syncthreads();
    uint64_t nbytes = 0;
    if (xt == 0 && tid < kk) {
        const int64_t j = static_cast<int64_t>(v.e) + tid;
        for (int t = 0; t < tn; ++t) {
            uint64_t x = T[t * kBlockK + tid];
            if ((j >> 6) == w0 + t) x &= ~(1ull << (j & 63));
            v.Sw[static_cast<int64_t>(w0 + t) * v.npad + j] = x;
        }
    }
    const int64_t idx = xt * blockDim.x + tid;
    if (idx < J) {
        const int64_t j = v.ulist[idx];
        const uint64_t b = v.Bc[j];
        for (int t0 = 0; t0 < tn; t0 += kW) {
            uint64_t xs[kW];
#pragma unroll
            for (int u = 0; u < kW; ++u) xs[u] = 0;
#pragma unroll 4
            for (int i = 0; i < kBlockK; ++i) {
                const uint64_t m = 0ull - ((b >> i) & 1ull);
#pragma unroll
                for (int u = 0; u < kW; ++u) xs[u] ^= T[(t0 + u) * kBlockK + i] & m;
            }
            uint64_t sv[kW];
#pragma unroll
            for (int u = 0; u < kW; ++u) {
                if (t0 + u >= tn) xs[u] = 0;
                sv[u] = xs[u] ? v.Sw[static_cast<int64_t>(w0 + t0 + u) * v.npad + j] : 0ull;
            }
#pragma unroll
            for (int u = 0; u < kW; ++u)
                if (xs[u]) {
                    v.Sw[static_cast<int64_t>(w0 + t0 + u) * v.npad + j] = sv[u] ^ xs[u];
                    nbytes += 16;
                }
        }
    }
    for (int o = kWave / 2; o > 0; o >>= 1) nbytes += __shfl_down(nbytes, o);
    if ((tid & (kWave - 1)) == 0 && nbytes) add_bytes(upd_bytes, nbytes);
}

// The XOR update of one block over (column tile, word tile) pairs, dealt to a fixed grid: column tile
// 0 is the block's own 64 columns, which take their final vectors; tiles 1.. are the listed later
// columns (ulist, 256 each), which take S_j ^= sum_i b_ji S_{e+i}.  Per pair the workgroup builds
// four-Russians tables of kTile words in LDS (Tab[t][g][x]: the XOR of the final vectors 4g+k, k in
// x, at word w0+t), so a word costs 16 lookups.  A tile whose final vectors are all zero changes no
// later column.  Columns whose coefficients are zero are never visited.
__device__ __forceinline__ void xor_body(const BlkView &v, uint64_t *lds, unsigned long long *upd_bytes) {
    static_assert(kTile * 16 == 256, "one table per thread of the workgroup");
    constexpr int kW = 8;
    const int tid = threadIdx.x;
    const int kk = v.kk;
    const int32_t wl = (v.e + kk + kWave - 1) / kWave;
    const int64_t ytiles = (wl + kTile - 1) / kTile;
    const uint32_t J = *v.ucount;
    const int64_t tiles = (1 + (static_cast<int64_t>(J) + blockDim.x - 1) / blockDim.x) * ytiles;
    uint64_t nbytes = 0;
    uint64_t *Tab = lds;  // [kTile][16][16]
    for (int64_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const int64_t ct = tile / ytiles;
        const int32_t w0 = static_cast<int32_t>(tile % ytiles) * kTile;
        const int tn = wl - w0 < kTile ? wl - w0 : kTile;
        int64_t j = -1;
        uint64_t b = 0;
        if (ct == 0) {
            if (tid < kk) j = static_cast<int64_t>(v.e) + tid;
        } else {
            const int64_t idx = (ct - 1) * blockDim.x + tid;
            if (idx < static_cast<int64_t>(J)) {
                j = v.ulist[idx];
                b = v.Bc[j];
            }
        }
        __syncthreads();  // the previous pair's tables are no longer read
        bool tile_nz;
        {
            const int t = tid >> 4, g = tid & 15;
            uint64_t r[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int i = 4 * g + k;
                r[k] = (t < tn && i < kk) ? v.Sf_blk[static_cast<int64_t>(w0 + t) * kBlockK + i] : 0ull;
            }
            tile_nz = (r[0] | r[1] | r[2] | r[3]) != 0;
            uint64_t e[16];
            e[0] = 0;
#pragma unroll
            for (int x = 1; x < 16; ++x) e[x] = e[x & (x - 1)] ^ r[__builtin_ctz(x)];
            uint64_t *row = Tab + (t * 16 + g) * 16;
#pragma unroll
            for (int x = 0; x < 16; ++x) row[x] = e[x];
        }
        __syncthreads();
        tile_nz = __syncthreads_or(tile_nz) != 0;
        if (ct == 0) {
            if (j >= 0) {
                const int i = static_cast<int>(j - v.e);
                for (int t = 0; t < tn; ++t) {
                    uint64_t x = Tab[(t * 16 + (i >> 2)) * 16 + (1 << (i & 3))];
                    if ((j >> 6) == w0 + t) x &= ~(1ull << (j & 63));
                    v.Sw[static_cast<int64_t>(w0 + t) * v.npad + j] = x;
                }
            }
        } else if (j >= 0 && tile_nz && b) {
            for (int t0 = 0; t0 < tn; t0 += kW) {
                uint64_t xs[kW];
#pragma unroll
                for (int u = 0; u < kW; ++u) xs[u] = 0;
#pragma unroll 2
                for (int g = 0; g < 16; ++g) {
                    const int nib = static_cast<int>((b >> (4 * g)) & 15ull);
#pragma unroll
                    for (int u = 0; u < kW; ++u) xs[u] ^= Tab[((t0 + u) * 16 + g) * 16 + nib];
                }
                uint64_t sv[kW];
#pragma unroll
                for (int u = 0; u < kW; ++u) {
                    if (t0 + u >= tn) xs[u] = 0;
                    sv[u] = xs[u] ? v.Sw[static_cast<int64_t>(w0 + t0 + u) * v.npad + j] : 0ull;
                }
#pragma unroll
                for (int u = 0; u < kW; ++u)
                    if (xs[u]) {
                        v.Sw[static_cast<int64_t>(w0 + t0 + u) * v.npad + j] = sv[u] ^ xs[u];
                        nbytes += 16;
                    }
            }
        }
    }
    for (int o = kWave / 2; o > 0; o >>= 1) nbytes += __shfl_down(nbytes, o);
    if ((tid & (kWave - 1)) == 0 && nbytes) add_bytes(upd_bytes, nbytes);
}

__global__ __launch_bounds__(256) void k_block_xor(BlockArgs a) {
    extern __shared__ uint64_t lds[];
    xor_body(view_of(a), lds, a.upd_bytes);
}

// ---------------------------------------------------------------------------------------------
// Batched blocked form: many independent problems (the biconnected components of one graph) in one
// ragged arena, advanced block by block together.  Step s runs block s (phases 64s .. 64s+63) of
// every problem that has that many phases, with one launch per kernel for all of them.  The
// problems are numbered by phase count, descending, so those still running at step s are a prefix.
// Tree positions keep their per-tree numbering; coordinates and ranks are shifted into disjoint
// ranges (coordinate base cb, rank base rb, both multiples of 64) so that Tt, Ct, Bc, Qr and the
// candidate bitmap are shared arrays.  A problem's candidates are odd only for its own support
// vectors, so running them side by side changes no problem's result.
// ---------------------------------------------------------------------------------------------

constexpr int kXcds = 8;  // accelerator complex dies of an MI355X; workgroups are dealt to them round-robin

struct CompInfo {
    int64_t cb;     // first coordinate (Tt, Ct, Bc, sel)
    int64_t rb;     // first rank (Qr, candidate bitmap, candidate arrays)
    int64_t sb;     // first word of the support matrix [W][npad]
    int64_t cblk;   // first word of C_blk, Sf_blk, D_blk, pair_i, pair_c ([64][W])
    int64_t pw;     // first entry of pw_list
    int64_t pwoff;  // first entry of pw_off
    int64_t store;  // first word of C_store ([N][W])
    int64_t npad;
    int64_t top;    // slot of the top-level basis
    int32_t N, W, levels, pad_;
};

struct ChunkInfo {  // first-level basis: ranks [r0, r0 + 64 * words) into slot
    int64_t r0, slot;
    int32_t comp, words;
};

struct MergeGroup {  // `members` bases from slot src on into slot dst
    int64_t src, dst;
    int32_t comp, members;
};

struct BatchArgs {
    const int64_t *tree_off;
    const uint32_t *parent_pos;
    const int32_t *pos_coord;  // shifted by the problem's cb
    const uint32_t *pos_pre, *pos_end;
    const ItemMeta *big_list, *small_list;  // work lists, problem-major, each problem's trees largest first
    uint64_t *scratch;
    int64_t huge_words;
    const int64_t *rc_off;
    const uint32_t *rc_pu, *rc_pv, *rc_rank;  // ranks shifted by rb
    const int32_t *rc_coord;
    const uint32_t *cand_root, *cand_pu, *cand_pv;  // indexed by shifted rank
    const int32_t *cand_coord;
    const uint32_t *root_comp;
    const CompInfo *comp;
    const ChunkInfo *chunks;
    const MergeGroup *groups;  // this launch's level
    const int32_t *pairs_cum;  // [problems + 1] workgroups of k_batch_pairs before each problem
    const int32_t *dot_cum;    // this step's [active + 1]
    const int32_t *xor_cum;    // this step's [active + 1]
    uint64_t *Sw, *Tt, *Qr, *qbits, *sum_q;
    uint32_t *sum_cnt, *sum_rank, *sel;
    uint64_t *blk_q, *blk_m, *blk_g, *C_blk, *Sf_blk, *C_store, *Ct, *Bc, *pair_c;
    uint32_t *D_blk, *Dcnt_blk, *pw_list, *pw_off, *pw_cnt, *pair_i;
    uint32_t *ulist;   // per problem at its coordinate base: the later columns that take this step's update
    uint32_t *ucount;  // [problems] their number (cleared by k_batch_transpose)
    uint32_t *status;
    unsigned long long *upd_bytes;
    int64_t K;  // rank bound
    int32_t step, active, n_items, big_words, small_words, xor_wt;
    int32_t x_big_off[kXcds], x_big_n[kXcds];      // per XCD section of the big-tree list: start, running entries
    int32_t x_small_off[kXcds], x_small_n[kXcds];  // the same for the small-tree list
};

__device__ __forceinline__ int32_t batch_kk(int32_t N, int32_t step) {
    const int32_t left = N - step * kBlockK;
    return left < kBlockK ? left : kBlockK;
}

__device__ __forceinline__ BlkView batch_view(const BatchArgs &a, int g) {
    const CompInfo &c = a.comp[g];
    BlkView v;
    v.Sw = a.Sw + uniform64(c.sb);
    v.npad = uniform64(c.npad);
    v.cb = uniform64(c.cb);
    v.N = static_cast<int32_t>(uniform32(static_cast<uint32_t>(c.N)));
    v.W = static_cast<int32_t>(uniform32(static_cast<uint32_t>(c.W)));
    v.e = a.step * kBlockK;
    v.kk = batch_kk(v.N, a.step);
    v.sel = a.sel + v.cb;
    v.blk_q = a.blk_q + static_cast<int64_t>(g) * kBlockK;
    v.blk_m = a.blk_m + static_cast<int64_t>(g) * kBlockK;
    v.blk_g = a.blk_g + static_cast<int64_t>(g) * kBlockK;
    const int64_t cblk = uniform64(c.cblk);
    v.C_blk = a.C_blk + cblk;
    v.Sf_blk = a.Sf_blk + cblk;
    v.D_blk = a.D_blk + cblk;
    v.pair_i = a.pair_i + cblk;
    v.pair_c = a.pair_c + cblk;
    v.ulist = a.ulist + v.cb;
    v.ucount = a.ucount + g;
    v.Dcnt_blk = a.Dcnt_blk + static_cast<int64_t>(g) * kBlockK;
    v.C_store = a.C_store ? a.C_store + uniform64(c.store) : nullptr;
    v.Tt = a.Tt + v.cb;
    v.Ct = a.Ct + v.cb;
    v.Bc = a.Bc + v.cb;
    v.pw_list = a.pw_list + uniform64(c.pw);
    v.pw_off = a.pw_off + uniform64(c.pwoff);
    v.pw_cnt = a.pw_cnt + g;
    return v;
}

// Largest g in [0, n) with cum[g] <= x (cum ascending from cum[0] = 0, x < cum[n]): the segment
// that holds item x.  Empty segments are skipped because the largest such g is taken.
__device__ __forceinline__ int find_segment(const int32_t *cum, int n, int32_t x) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (static_cast<int32_t>(uniform32(static_cast<uint32_t>(cum[mid]))) <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Wave x: word x % (step + 1) of problem x / (step + 1) (every running problem has step + 1 words
// below the block's end).
__global__ __launch_bounds__(256) void k_batch_transpose(BatchArgs a) {
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t x = static_cast<int64_t>(blockIdx.x) * 4 + wave_id();
    const int64_t wl = a.step + 1;
    if (x >= static_cast<int64_t>(a.active) * wl) return;
    const int g = static_cast<int>(x / wl);
    if (x == static_cast<int64_t>(g) * wl && lane == 0) a.ucount[g] = 0;  // k_batch_dot lists problem g's columns
    transpose_body(batch_view(a, g), x - g * wl, lane);
}

// Workgroups run on the XCD numbered blockIdx.x mod kXcds (round-robin dispatch); problem g's trees
// are listed in section g mod kXcds, so that each problem's Tt words are gathered through one XCD's L2.
// Six waves per SIMD (80 VGPRs): built for seven as k_block_q, the 2000-component batch's tree pass took
// 2240 us per step instead of 2192.
__global__ __launch_bounds__(kQWaves * kWave, 5) void k_batch_q(BatchArgs a) {
    const uint32_t nw = blockDim.x / kWave;
    // blocks x, x + kXcds, ... serve section x (the grid is a multiple of kXcds)
    const uint32_t x = blockIdx.x % kXcds, local = blockIdx.x / kXcds, nloc = gridDim.x / kXcds;
    const uint32_t bo = static_cast<uint32_t>(a.x_big_off[x]), so = static_cast<uint32_t>(a.x_small_off[x]);
    tree_pass(a, a.big_list + bo, local, static_cast<uint32_t>(a.x_big_n[x]), nloc,
              a.small_list + so, local * nw + static_cast<uint32_t>(wave_id()), static_cast<uint32_t>(a.x_small_n[x]),
              nloc * nw, [&](uint32_t r) {
                  const uint32_t g = uniform32(a.root_comp[r]);
                  return batch_kk(static_cast<int32_t>(uniform32(static_cast<uint32_t>(a.comp[g].N))), a.step);
              });
}

__device__ __forceinline__ void batch_finalize(const BatchArgs &a, int g, int lane) {
    const CompInfo &c = a.comp[g];
    const int32_t kk = batch_kk(static_cast<int32_t>(uniform32(static_cast<uint32_t>(c.N))), a.step);
    finalize_core(kk, a.sum_cnt, a.sum_rank, a.sum_q, uniform64(c.top),
                  a.sel + uniform64(c.cb) + static_cast<int64_t>(a.step) * kBlockK,
                  a.blk_q + static_cast<int64_t>(g) * kBlockK, a.blk_m + static_cast<int64_t>(g) * kBlockK,
                  a.blk_g + static_cast<int64_t>(g) * kBlockK, a.status, lane);
}

// Wave x: first-level chunk x (of the running problems' chunks, problem-major).
__global__ __launch_bounds__(256) void k_batch_basis(BatchArgs a) {
    __shared__ uint64_t s_words[4][kWave];
    __shared__ uint32_t s_offs[4][kWave];
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = wave_id();
    const int64_t x = static_cast<int64_t>(blockIdx.x) * 4 + wv;
    if (x >= a.n_items) return;
    const ChunkInfo &ch = a.chunks[x];
    const int g = static_cast<int>(uniform32(static_cast<uint32_t>(ch.comp)));
    const CompInfo &c = a.comp[g];
    const int32_t kk = batch_kk(static_cast<int32_t>(uniform32(static_cast<uint32_t>(c.N))), a.step);
    basis_chunk<1>(a.qbits, a.Qr, uniform64(ch.r0), static_cast<int>(uniform32(static_cast<uint32_t>(ch.words))),
                uniform64(ch.slot), kk, a.sum_cnt, a.sum_rank, a.sum_q, s_words[wv], s_offs[wv], lane);
    if (uniform32(static_cast<uint32_t>(c.levels)) == 1u) {
        __threadfence();
        batch_finalize(a, g, lane);
    }
}

// Wave x: merge group x of level L (of the running problems that have a level L+1); a problem's top
// group also finalises its block.
__global__ __launch_bounds__(kMergeWaves * kWave) void k_batch_merge(BatchArgs a, int L) {
    __shared__ uint64_t s_tab[kMergeTab];
    __shared__ uint64_t s_stq[kMergeEntries];
    __shared__ uint32_t s_str[kMergeEntries];
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = wave_id();
    const int64_t x = blockIdx.x;
    if (x >= a.n_items) return;
    const MergeGroup &mg = a.groups[x];
    const int g = static_cast<int>(uniform32(static_cast<uint32_t>(mg.comp)));
    const CompInfo &c = a.comp[g];
    const int32_t kk = batch_kk(static_cast<int32_t>(uniform32(static_cast<uint32_t>(c.N))), a.step);
    if (!merge_group(uniform64(mg.src), static_cast<int>(uniform32(static_cast<uint32_t>(mg.members))),
                     uniform64(mg.dst), kk, a.sum_cnt, a.sum_rank, a.sum_q, s_tab, s_stq, s_str, lane, wv))
        return;
    if (static_cast<int>(uniform32(static_cast<uint32_t>(c.levels))) == L + 2) {
        __threadfence();
        batch_finalize(a, g, lane);
    }
}

// Workgroup 64 g + i: phase 64*step + i of problem g.
__global__ __launch_bounds__(256) void k_batch_cycles(BatchArgs a) {
    extern __shared__ uint64_t lds[];
    __shared__ uint32_t cnt;
    const int g = static_cast<int>(blockIdx.x / kBlockK), i = static_cast<int>(blockIdx.x % kBlockK);
    const BlkView v = batch_view(a, g);
    if (i >= v.kk) return;
    cycles_body(a, v, i, lds, &cnt);
}

// Wave x: word x % (step + 1) of problem x / (step + 1)'s final support vectors.
__global__ __launch_bounds__(256) void k_batch_final(BatchArgs a) {
    const int64_t x = static_cast<int64_t>(blockIdx.x) * 4 + wave_id();
    const int64_t wl = a.step + 1;
    if (x >= static_cast<int64_t>(a.active) * wl) return;
    const int g = static_cast<int>(x / wl);
    final_body(batch_view(a, g), x - g * wl, threadIdx.x & (kWave - 1));
}

// Problem g owns workgroups pairs_cum[g] .. pairs_cum[g+1]-1: the first lists its words, the others
// transpose its cycle vectors.
__global__ __launch_bounds__(256) void k_batch_pairs(BatchArgs a) {
    extern __shared__ uint64_t lds[];
    __shared__ uint32_t nlist, npairs;
    const int32_t x = static_cast<int32_t>(blockIdx.x);
    const int g = find_segment(a.pairs_cum, a.active, x);
    const BlkView v = batch_view(a, g);
    const int32_t b = x - a.pairs_cum[g];
    if (b > 0) {
        const int64_t w = static_cast<int64_t>(b - 1) * 4 + wave_id();
        if (w < v.W) pairs_transpose(v, w, threadIdx.x & (kWave - 1));
        return;
    }
    pairs_list(v, lds, &nlist, &npairs);
}

__global__ __launch_bounds__(kDotWaves * kWave) void k_batch_dot(BatchArgs a) {
    extern __shared__ uint64_t lds[];
    __shared__ uint32_t seg_end;
    const int32_t x = static_cast<int32_t>(blockIdx.x);
    const int g = find_segment(a.dot_cum, a.active, x);
    dot_body(batch_view(a, g), x - a.dot_cum[g], lds, &seg_end, a.upd_bytes);
}

// Problem g's tiles: ceil((N - 64 step) / 256) column tiles times xor_wt word tiles, columns fastest.
__global__ __launch_bounds__(256) void k_batch_xor(BatchArgs a) {
    extern __shared__ uint64_t lds[];
    const int32_t x = static_cast<int32_t>(blockIdx.x);
    const int g = find_segment(a.xor_cum, a.active, x);
    const BlkView v = batch_view(a, g);
    const int32_t t = x - a.xor_cum[g];
    const int32_t ct = (v.N - v.e + 255) / 256;
    const int32_t yt = static_cast<int32_t>(uniform32(static_cast<uint32_t>(t / ct)));
    xor_direct(v, t - yt * ct, yt, lds, a.upd_bytes);
}

// ---------------------------------------------------------------------------------------------
// Device properties shared by both forms
// ---------------------------------------------------------------------------------------------

struct DeviceProps {
    bool init = false;
    int device = 0;
    hipDeviceProp_t prop{};
};
DeviceProps g_dev;

int ensure_device() {
    if (g_dev.init) return MCB_OK;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return fail(MCB_E_DEVICE, "no HIP device");
    HIP_TRY(hipGetDevice(&g_dev.device));
    HIP_TRY(hipGetDeviceProperties(&g_dev.prop, g_dev.device));
    g_dev.init = true;
    return MCB_OK;
}

template <typename T>
int dev_alloc(T **p, size_t count, int64_t &total) {
    *p = nullptr;
    if (count == 0) count = 1;
    HIP_TRY(hipMalloc(reinterpret_cast<void **>(p), count * sizeof(T)));
    total += static_cast<int64_t>(count * sizeof(T));
    return MCB_OK;
}

template <typename T>
int upload(T **p, const T *src, size_t count, int64_t &total) {
    int rc = dev_alloc(p, count, total);
    if (rc) return rc;
    if (count && src) HIP_TRY(hipMemcpy(*p, src, count * sizeof(T), hipMemcpyHostToDevice));
    return MCB_OK;
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// Reference-compatible form
// ---------------------------------------------------------------------------------------------

struct mcb_gpu {
    int num_edges = 0, num_non_tree = 0, size_vector = 0, nodes = 0, fvs = 0, chunk = 0;
    int num_chunks = 0, nslots = 0, nstreams = 0;
    bool all_resident = false;
    int32_t *d_nt = nullptr;
    uint64_t *d_S = nullptr;
    std::vector<int32_t *> d_eo;
    std::vector<uint32_t *> d_cols, d_out;
    std::vector<hipStream_t> streams;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool initialized = false;
    int64_t bytes = 0;
};

namespace {

int compat_roots_in_chunk(const mcb_gpu *g, int i) { return std::min(g->chunk, g->fvs - i * g->chunk); }

// The kernels index LDS by parent position and the coordinate map by edge offset, so both are
// range-checked on the host before a chunk is used.
int compat_check_layout(const mcb_gpu *g, const uint32_t *cols, const int32_t *eo, int roots) {
    for (int s = 0; s < roots; ++s) {
        const uint32_t *c = cols + static_cast<size_t>(s) * g->nodes;
        const int32_t *e = eo + static_cast<size_t>(s) * g->nodes;
        if (c[0] != kNoParent) return fail(MCB_E_INVALID, "tree_cols: position 0 must be the root (parent -1)");
        for (int i = 1; i < g->nodes; ++i)
            if (c[i] >= static_cast<uint32_t>(i))
                return fail(MCB_E_INVALID, "tree_cols: a parent position must precede its child");
        for (int i = 0; i < g->nodes; ++i)
            if (e[i] < -1 || e[i] >= g->num_edges) return fail(MCB_E_INVALID, "edge_offset out of range");
    }
    return MCB_OK;
}

int compat_launch(mcb_gpu *g, int slot, int roots, hipStream_t st) {
    const int p_words = (g->nodes + 63) / 64;
    int wpb = 4;
    while (wpb > 1 && wpb * p_words * 8 > kLdsLimit) --wpb;
    if (p_words * 8 > kLdsLimit) return fail(MCB_E_INVALID, "tree too large for the LDS parity buffer");
    const int blocks = (roots + wpb - 1) / wpb;
    hipLaunchKernelGGL(k_compat_parity, dim3(blocks), dim3(wpb * kWave), wpb * p_words * 8, st, g->d_nt,
                       g->d_eo[slot], g->d_cols[slot], g->d_out[slot], g->d_S, g->nodes, roots, p_words);
    HIP_TRY(hipGetLastError());
    return MCB_OK;
}

}  // namespace

extern "C" {

const char *mcb_hip_last_error(void) { return g_err.c_str(); }

int mcb_hip_device_count(void) {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) return 0;
    return c;
}

int mcb_hip_device_info(char *name, int name_len, int *compute_units, int64_t *total_mem) {
    int rc = ensure_device();
    if (rc) return rc;
    if (name && name_len > 0) {
        if (g_dev.prop.name[0])
            std::snprintf(name, static_cast<size_t>(name_len), "%s (%s)", g_dev.prop.name, g_dev.prop.gcnArchName);
        else
            std::snprintf(name, static_cast<size_t>(name_len), "%s, %d CUs", g_dev.prop.gcnArchName,
                          g_dev.prop.multiProcessorCount);
    }
    if (compute_units) *compute_units = g_dev.prop.multiProcessorCount;
    if (total_mem) *total_mem = static_cast<int64_t>(g_dev.prop.totalGlobalMem);
    return MCB_OK;
}

void init_cuda(void) {
    HIP_OR_EXIT(hipGetDevice(&g_dev.device));
    HIP_OR_EXIT(hipGetDeviceProperties(&g_dev.prop, g_dev.device));
    g_dev.init = true;
}

size_t configure_grid(int start, int end) {
    if (!g_dev.init) init_cuda();
    const int cus = g_dev.prop.multiProcessorCount;
    return (end - start) < cus ? static_cast<size_t>(end - start) : static_cast<size_t>(cus);
}

unsigned *allocate_pinned_memory(int chunk, int nodes) {
    unsigned *p = nullptr;
    HIP_OR_EXIT(hipHostMalloc(reinterpret_cast<void **>(&p),
                              sizeof(unsigned) * static_cast<size_t>(chunk) * static_cast<size_t>(nodes),
                              hipHostMallocDefault));
    return p;
}

void free_pinned_memory(unsigned *p) { HIP_OR_EXIT(hipHostFree(p)); }

size_t calculate_chunk_size(size_t num_nodes, size_t num_edges, size_t size_vector, size_t nstream) {
    if (!g_dev.init) init_cuda();
    const size_t global_bytes = g_dev.prop.totalGlobalMem;
    const size_t static_bytes = 4 * num_edges + 8 * size_vector;
    if (static_bytes >= global_bytes || nstream == 0) return 0;
    const size_t elems = (global_bytes - static_bytes) / 4;
    return elems / (nstream * (num_nodes * 4 + 1));
}

int mcb_gpu_create(int num_edges, int num_non_tree_edges, int size_vector, int original_nodes, int fvs_size,
                   int chunk_size, int nstreams, mcb_gpu **out) {
    return mcb_gpu_create_slots(num_edges, num_non_tree_edges, size_vector, original_nodes, fvs_size, chunk_size,
                                nstreams, 0, out);
}

int mcb_gpu_create_slots(int num_edges, int num_non_tree_edges, int size_vector, int original_nodes, int fvs_size,
                         int chunk_size, int nstreams, int max_slots, mcb_gpu **out) {
    if (!out || num_edges <= 0 || num_non_tree_edges <= 0 || size_vector <= 0 || original_nodes <= 0 ||
        fvs_size <= 0 || chunk_size <= 0 || nstreams <= 0 || max_slots < 0)
        return fail(MCB_E_INVALID, "invalid arguments");
    *out = nullptr;
    int rc = ensure_device();
    if (rc) return rc;
    mcb_gpu *g = new (std::nothrow) mcb_gpu();
    if (!g) return fail(MCB_E_NOMEM, "out of memory");
    g->num_edges = num_edges;
    g->num_non_tree = num_non_tree_edges;
    g->size_vector = size_vector;
    g->nodes = original_nodes;
    g->fvs = fvs_size;
    g->chunk = chunk_size;
    g->num_chunks = (fvs_size + chunk_size - 1) / chunk_size;
    const size_t per_slot = static_cast<size_t>(chunk_size) * original_nodes * 12;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = 0;
    const size_t reserve = static_cast<size_t>(1) << 30;
    g->all_resident = static_cast<size_t>(g->num_chunks) * per_slot + reserve < free_b &&
                      (max_slots == 0 || g->num_chunks <= max_slots);
    g->nslots = g->all_resident ? g->num_chunks : std::min(max_slots > 0 ? max_slots : 32, g->num_chunks);
    g->nstreams = std::min(32, std::max(1, std::min(nstreams, g->num_chunks)));
    auto cleanup = [&](int code) {
        mcb_gpu_destroy(g);
        return code;
    };
    if ((rc = dev_alloc(&g->d_nt, num_edges, g->bytes))) return cleanup(rc);
    if ((rc = dev_alloc(&g->d_S, size_vector, g->bytes))) return cleanup(rc);
    g->d_eo.assign(g->nslots, nullptr);
    g->d_cols.assign(g->nslots, nullptr);
    g->d_out.assign(g->nslots, nullptr);
    const size_t cn = static_cast<size_t>(chunk_size) * original_nodes;
    for (int i = 0; i < g->nslots; ++i) {
        if ((rc = dev_alloc(&g->d_eo[i], cn, g->bytes))) return cleanup(rc);
        if ((rc = dev_alloc(&g->d_cols[i], cn, g->bytes))) return cleanup(rc);
        if ((rc = dev_alloc(&g->d_out[i], cn, g->bytes))) return cleanup(rc);
    }
    g->streams.assign(g->nstreams, nullptr);
    for (int i = 0; i < g->nstreams; ++i)
        if (hipStreamCreate(&g->streams[i]) != hipSuccess) return cleanup(fail(MCB_E_DEVICE, "hipStreamCreate"));
    if (hipEventCreate(&g->ev0) != hipSuccess || hipEventCreate(&g->ev1) != hipSuccess)
        return cleanup(fail(MCB_E_DEVICE, "hipEventCreate"));
    *out = g;
    return MCB_OK;
}

int mcb_gpu_initialize_memory(mcb_gpu *g, const int32_t *nt, const uint32_t *const *tree_rows,
                              const uint32_t *const *tree_cols, const int32_t *const *edge_offset, int num_rows) {
    (void)tree_rows;
    if (!g || !nt || !tree_cols || !edge_offset || num_rows != g->num_chunks)
        return fail(MCB_E_INVALID, "invalid arguments");
    if (static_cast<int64_t>(g->num_non_tree) > 64 * static_cast<int64_t>(g->size_vector))
        return fail(MCB_E_INVALID, "size_vector too small for num_non_tree_edges");
    for (int i = 0; i < g->num_edges; ++i)
        if (nt[i] < -1 || nt[i] >= g->num_non_tree) return fail(MCB_E_INVALID, "non-tree map entry out of range");
    HIP_TRY(hipMemcpy(g->d_nt, nt, sizeof(int32_t) * g->num_edges, hipMemcpyHostToDevice));
    for (int i = 0; i < g->num_chunks; ++i) {
        const int roots = compat_roots_in_chunk(g, i);
        int rc = compat_check_layout(g, tree_cols[i], edge_offset[i], roots);
        if (rc) return rc;
        if (i < g->nslots && g->all_resident) {
            const size_t bytes = sizeof(int32_t) * static_cast<size_t>(roots) * g->nodes;
            HIP_TRY(hipMemcpy(g->d_eo[i], edge_offset[i], bytes, hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(g->d_cols[i], tree_cols[i], bytes, hipMemcpyHostToDevice));
        }
    }
    g->initialized = true;
    return MCB_OK;
}

int mcb_gpu_copy_support_vector(mcb_gpu *g, const uint64_t *elements, float *ms) {
    if (!g || !elements) return fail(MCB_E_INVALID, "invalid arguments");
    HIP_TRY(hipEventRecord(g->ev0, nullptr));
    HIP_TRY(hipMemcpy(g->d_S, elements, sizeof(uint64_t) * g->size_vector, hipMemcpyHostToDevice));
    HIP_TRY(hipEventRecord(g->ev1, nullptr));
    HIP_TRY(hipEventSynchronize(g->ev1));
    if (ms) HIP_TRY(hipEventElapsedTime(ms, g->ev0, g->ev1));
    return MCB_OK;
}

int mcb_gpu_process_shortest_path(mcb_gpu *g, const uint32_t *const *tree_rows, const uint32_t *const *tree_cols,
                                  const int32_t *const *edge_offset, uint32_t *const *precompute_value, int num_rows,
                                  int multiple_transfer, float *ms) {
    (void)tree_rows;
    if (!g || !precompute_value || num_rows != g->num_chunks) return fail(MCB_E_INVALID, "invalid arguments");
    if (!g->initialized) return fail(MCB_E_STATE, "initialize_memory has not been called");
    if (multiple_transfer && (!tree_cols || !edge_offset)) return fail(MCB_E_INVALID, "trees needed for transfer");
    if (!multiple_transfer && !g->all_resident)
        return fail(MCB_E_STATE, "trees do not fit in device memory; use multiple_transfer");
    HIP_TRY(hipEventRecord(g->ev0, nullptr));
    for (int i = 0; i < g->num_chunks; ++i) {
        const int slot = g->all_resident ? i : i % g->nslots;
        hipStream_t st = g->streams[i % g->nstreams];
        const int roots = compat_roots_in_chunk(g, i);
        const size_t bytes = sizeof(uint32_t) * static_cast<size_t>(roots) * g->nodes;
        if (multiple_transfer) {
            int rc = compat_check_layout(g, tree_cols[i], edge_offset[i], roots);
            if (rc) return rc;
            HIP_TRY(hipMemcpyAsync(g->d_eo[slot], edge_offset[i], bytes, hipMemcpyHostToDevice, st));
            HIP_TRY(hipMemcpyAsync(g->d_cols[slot], tree_cols[i], bytes, hipMemcpyHostToDevice, st));
        }
        int rc = compat_launch(g, slot, roots, st);
        if (rc) return rc;
        HIP_TRY(hipMemcpyAsync(precompute_value[i], g->d_out[slot], bytes, hipMemcpyDeviceToHost, st));
        if (!g->all_resident && (i + 1) % g->nslots == 0) HIP_TRY(hipDeviceSynchronize());
    }
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipEventRecord(g->ev1, nullptr));
    HIP_TRY(hipEventSynchronize(g->ev1));
    if (ms) HIP_TRY(hipEventElapsedTime(ms, g->ev0, g->ev1));
    return MCB_OK;
}

int mcb_gpu_fetch(mcb_gpu *g, uint32_t *const *precompute_value, int num_rows, float *ms) {
    if (!g || !precompute_value || num_rows != g->num_chunks) return fail(MCB_E_INVALID, "invalid arguments");
    HIP_TRY(hipEventRecord(g->ev0, nullptr));
    for (int i = 0; i < g->num_chunks && i < g->nslots; ++i) {
        const size_t bytes = sizeof(uint32_t) * static_cast<size_t>(compat_roots_in_chunk(g, i)) * g->nodes;
        HIP_TRY(hipMemcpy(precompute_value[i], g->d_out[i], bytes, hipMemcpyDeviceToHost));
    }
    HIP_TRY(hipEventRecord(g->ev1, nullptr));
    HIP_TRY(hipEventSynchronize(g->ev1));
    if (ms) HIP_TRY(hipEventElapsedTime(ms, g->ev0, g->ev1));
    return MCB_OK;
}

int mcb_gpu_calculate_memory(mcb_gpu *g, double *total_mb, double *static_mb, double *variable_mb) {
    if (!g) return fail(MCB_E_INVALID, "invalid arguments");
    const double mb = 1024.0 * 1024.0;
    const double st = (4.0 * g->num_edges + 8.0 * g->size_vector) / mb;
    const double var = static_cast<double>(g->bytes) / mb - st;
    if (total_mb) *total_mb = st + var;
    if (static_mb) *static_mb = st;
    if (variable_mb) *variable_mb = var;
    return MCB_OK;
}

void mcb_gpu_destroy(mcb_gpu *g) {
    if (!g) return;
    (void)hipFree(g->d_nt);
    (void)hipFree(g->d_S);
    for (auto p : g->d_eo) (void)hipFree(p);
    for (auto p : g->d_cols) (void)hipFree(p);
    for (auto p : g->d_out) (void)hipFree(p);
    for (auto s : g->streams)
        if (s) (void)hipStreamDestroy(s);
    if (g->ev0) (void)hipEventDestroy(g->ev0);
    if (g->ev1) (void)hipEventDestroy(g->ev1);
    delete g;
}

}  // extern "C"

// ---------------------------------------------------------------------------------------------
// Native form
// ---------------------------------------------------------------------------------------------

namespace {

// Launch plan of the blocked tree pass for nw waves per workgroup.  Trees with more than small_max
// positions are shared by the whole workgroup, the others take one wave; small_max is the largest
// multiple of 64 (at most 1024) whose nw one-wave slices fit the LDS the largest tree needs anyway.
// The roots come in ranges (one per problem of a batch, one for a single problem), each range's trees
// largest first.  Range g is listed in section g mod nsplit (sections one after the other);
// big_cum[x][g] / small_cum[x][g] count section x's entries from ranges before g.
// Entries are items: a root and a run of its candidates.  A big tree whose candidates exceed both
// kItemCands and the plan's mean candidates per workgroup is listed once per run of that size, so that
// a root holding a large share of all candidates spreads over several workgroups (each recomputes the
// path words).  Big trees whose path words do not fit the LDS next to their ready stamps keep them in
// scratch, huge_words words per workgroup.
constexpr int64_t kItemCands = 2048;

struct TreePlan {
    int32_t nw = 0, big_words = 0, small_words = 0, lds = 0, blocks = 0, n_big = 0, n_small = 0;
    int64_t huge_words = 0;
    ItemMeta *big = nullptr, *small = nullptr;  // the work lists (big items flagged narrow in npos)
    // one problem: big items of workgroup b are big[wg_big[b] .. wg_big[b+1]), small items of wave c
    // (workgroup c / nw) small[wave_small[c] .. wave_small[c+1]) (null for a batch: round-robin)
    int32_t *wg_big = nullptr, *wave_small = nullptr;
    uint64_t *scratch = nullptr;
    std::vector<int32_t> big_off, small_off;  // first entry of each section
    std::vector<std::vector<int32_t>> big_cum, small_cum;
    std::vector<int64_t> pos_cum, cand_cum;   // tree positions and candidates streamed, ranges before g
};

// The scratch is shared by a solver's plans (one runs at a time) and freed by its owner.
void free_plan(TreePlan &p) {
    void *ptrs[] = {p.big, p.small, p.wg_big, p.wave_small};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    p.big = p.small = nullptr;
    p.wg_big = p.wave_small = nullptr;
    p.scratch = nullptr;
}

}  // namespace

struct mcb_solver {
    mcb_solver_options opt{};
    int32_t F = 0, N = 0, W = 0, max_pos = 0, p_words_big = 0, p_words_small = 0, n_big = 0, n_small = 0;
    int64_t K = 0, Ptot = 0, npad = 0;
    int32_t phase = 0;
    // trees and candidates
    int64_t *tree_off = nullptr;
    uint32_t *parent_pos = nullptr, *pos_pre = nullptr, *pos_end = nullptr;
    int32_t *pos_coord = nullptr;
    uint32_t *big_list = nullptr, *small_list = nullptr, *root_min_rank = nullptr;
    int64_t *rc_off = nullptr;
    uint32_t *rc_pu = nullptr, *rc_pv = nullptr, *rc_rank = nullptr;
    int32_t *rc_coord = nullptr;
    uint32_t *cand_root = nullptr, *cand_pu = nullptr, *cand_pv = nullptr;
    int32_t *cand_coord = nullptr;
    // phase state: S_e contiguous in a ring of three, C_e and its word list in a ring of two
    uint64_t *Sw = nullptr, *Scur = nullptr, *C = nullptr, *C_store = nullptr;
    uint32_t *D = nullptr, *Dcount = nullptr, *best = nullptr, *done = nullptr, *status = nullptr;
    uint32_t *E = nullptr, *Ecount = nullptr;  // nonzero words of S_e, ring of four
    uint32_t *cdot = nullptr;                 // <C_e, S_{e-1}>, ring of two
    uint8_t *flag = nullptr;
    uint64_t *wave_stats = nullptr;  // 3 counters per wave of the phase kernel
    // launch shape
    int wpb = kBlockWaves, blocks = 1, lds_bytes = 0;
    hipStream_t stream = nullptr;
    int64_t bytes = 0;
    // blocked form (block_phases > 1)
    int32_t bk = 1;                                   // phases per block; 1 = one kernel per phase
    int32_t q_waves = 0, q_big_words = 0, q_small_words = 0, q_blocks = 0, q_lds = 0, n_qbig = 0, n_qsmall = 0;
    ItemMeta *q_big_list = nullptr, *q_small_list = nullptr;  // the chosen plan's lists
    int32_t *q_wg_big = nullptr, *q_wave_small = nullptr;
    std::vector<TreePlan> plans;
    int32_t plan = 0;
    uint64_t *q_scratch = nullptr;  // path words of trees beyond a plan's LDS cap
    hipGraph_t graph = nullptr;     // the whole loop, captured by mcb_solver_run_graph
    hipGraphExec_t graph_exec = nullptr;
    int64_t graph_positions = 0, graph_candidates = 0;  // tree pass work of one replay
    int32_t dot_wg = static_cast<int32_t>(kDotWorkgroups), xor_wg = static_cast<int32_t>(kXorWorkgroups);  // update grids
    hipEvent_t *tune_ev = nullptr;  // while the update grids are timed: events around each update launch
    size_t tune_n = 0, tune_cap = 0;
    uint64_t *Tt = nullptr, *Qr = nullptr, *qbits = nullptr, *sum_q = nullptr, *blk_q = nullptr, *blk_m = nullptr, *blk_g = nullptr;
    uint64_t *C_blk = nullptr, *Sf_blk = nullptr, *Ct = nullptr, *Bc = nullptr, *pair_c = nullptr;
    uint32_t *pw_list = nullptr, *pw_off = nullptr, *pw_cnt = nullptr, *pair_i = nullptr;
    uint32_t *sum_cnt = nullptr, *sum_rank = nullptr, *D_blk = nullptr, *Dcnt_blk = nullptr;
    uint32_t *arrive = nullptr;  // merge group arrivals of k_block_cascade
    uint32_t *ulist = nullptr, *ucount = nullptr;  // the block's updated later columns
    unsigned long long *upd_bytes = nullptr;
    int32_t levels = 0, group = 16, chunk_words = kWave;
    int64_t lvl_slot[kMaxLevels] = {}, lvl_n[kMaxLevels] = {};
    int64_t blk_positions = 0, blk_candidates = 0;  // work of the tree passes since reset
    // timing
    double kt_ms[MCB_K_COUNT] = {};
    int64_t kt_n[MCB_K_COUNT] = {};
    std::vector<hipEvent_t> ev_pool;
    std::vector<int> ev_kind;
    size_t ev_used = 0;
    bool timing_open = false;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;  // the open group's events, until launches take them
    bool timing_sample = true;  // this block (phase, step) records events
    int64_t timing_seq = 0;     // blocks (phases, steps) launched since reset
};

namespace {

template <typename S>
int harvest_events(S *s) {
    if (s->ev_used == 0) return MCB_OK;
    HIP_TRY(hipEventSynchronize(s->ev_pool[s->ev_used - 1]));
    for (size_t i = 0; i + 1 < s->ev_used; i += 2) {
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, s->ev_pool[i], s->ev_pool[i + 1]));
        s->kt_ms[s->ev_kind[i / 2]] += ms;
        s->kt_n[s->ev_kind[i / 2]] += 1;
    }
    s->ev_used = 0;
    return MCB_OK;
}

// Called once per block of phases (phase, step) before its launches: whether its kernels record events.
template <typename S>
void timing_sample(S *s) {
    const int64_t every = s->opt.timing_every > 1 ? s->opt.timing_every : 1;
    s->timing_sample = s->timing_seq++ % every == 0;
}

template <typename S>
int timed_begin(S *s, int kind) {
    s->timing_open = s->timing_sample && (s->opt.timing == 1 || (s->opt.timing == 2 && (kind == MCB_K_PARITY_SEARCH ||
                                                                                           kind == MCB_K_UPDATE)));
    if (!s->timing_open) return MCB_OK;
    if (s->ev_used + 2 > s->ev_pool.size()) {
        int rc = harvest_events(s);
        if (rc) return rc;
    }
    s->ev_kind[s->ev_used / 2] = kind;
    s->ev_start = s->ev_pool[s->ev_used++];
    s->ev_stop = s->ev_pool[s->ev_used++];
    return MCB_OK;
}

// Closes the group timed_begin opened.  Its launches carried its events; a group with no launch
// records them as markers.
template <typename S>
int timed_end(S *s) {
    if (!s->timing_open) return MCB_OK;
    s->timing_open = false;
    if (s->ev_start) HIP_TRY(hipEventRecord(s->ev_start, s->stream));
    if (s->ev_stop) HIP_TRY(hipEventRecord(s->ev_stop, s->stream));
    s->ev_start = s->ev_stop = nullptr;
    return MCB_OK;
}

// Launches a kernel on the solver's (batch's) stream.  In an open timing group the group's first
// launch carries the start event and the launch flagged `last` the stop event; the runtime stamps
// them with those kernels' own start and end, so the group's time holds no marker packets and no
// gap before its first kernel or after its last.  Launches without events go through the plain path
// (graph capture uses it).
template <typename S, typename... KArgs, typename... Args>
int launch_on(S *s, bool last, void (*kernel)(KArgs...), dim3 grid, dim3 block, size_t lds, Args... args) {
    hipEvent_t start = s->ev_start, stop = last ? s->ev_stop : nullptr;
    s->ev_start = nullptr;
    if (last) s->ev_stop = nullptr;
    if (start || stop)
        hipExtLaunchKernelGGL(kernel, grid, block, static_cast<uint32_t>(lds), s->stream, start, stop, 0u, args...);
    else
        hipLaunchKernelGGL(kernel, grid, block, lds, s->stream, args...);
    HIP_TRY(hipGetLastError());
    return MCB_OK;
}

PhaseArgs phase_args(const mcb_solver *s, int32_t e) {
    PhaseArgs a{};
    a.tree_off = s->tree_off;
    a.parent_pos = s->parent_pos;
    a.pos_coord = s->pos_coord;
    a.pos_pre = s->pos_pre;
    a.pos_end = s->pos_end;
    a.big_list = s->big_list;
    a.small_list = s->small_list;
    a.root_min_rank = s->root_min_rank;
    a.rc_off = s->rc_off;
    a.rc_pu = s->rc_pu;
    a.rc_pv = s->rc_pv;
    a.rc_rank = s->rc_rank;
    a.rc_coord = s->rc_coord;
    a.cand_root = s->cand_root;
    a.cand_pu = s->cand_pu;
    a.cand_pv = s->cand_pv;
    a.cand_coord = s->cand_coord;
    a.Sw = s->Sw;
    a.npad = s->npad;
    const int64_t W = s->W;
    a.S = s->Scur + (e % 4) * W;
    a.S_m1 = s->Scur + ((e + 3) % 4) * W;
    a.S_m2 = s->Scur + ((e + 2) % 4) * W;
    a.S_next = s->Scur + ((e + 1) % 4) * W;
    a.E_m2 = s->E + ((e + 2) % 4) * W;
    a.Ecount_m2 = s->Ecount + (e + 2) % 4;
    a.E_next = s->E + ((e + 1) % 4) * W;
    a.Ecount_next = s->Ecount + (e + 1) % 4;
    a.C_prev = s->C + ((e + 1) & 1) * W;
    a.D_prev = s->D + ((e + 1) & 1) * W;
    a.Dcount_prev = s->Dcount + ((e + 1) & 1);
    a.c_prev = s->cdot + ((e + 1) & 1);
    a.C = s->C + (e & 1) * W;
    a.D = s->D + (e & 1) * W;
    a.Dcount = s->Dcount + (e & 1);
    a.c_out = s->cdot + (e & 1);
    a.C_store = s->C_store ? s->C_store + static_cast<int64_t>(e) * W : nullptr;
    a.flag = s->flag;
    a.best = s->best + e;
    a.done = s->done + e;
    a.status = s->status;
    a.parity_out = nullptr;
    a.wave_stats = s->wave_stats;
    a.e = e;
    a.N = s->N;
    a.W = s->W;
    a.n_big = s->n_big;
    a.n_small = s->n_small;
    a.p_words_big = s->p_words_big;
    a.p_words_small = s->p_words_small;
    if (e >= 1 && e + 1 < s->N) a.upd_tiles = (s->N - (e + 1) + kWave - 1) / kWave;
    return a;
}

BlockArgs block_args(const mcb_solver *s, int32_t e, int32_t kk) {
    BlockArgs a{};
    a.tree_off = s->tree_off;
    a.parent_pos = s->parent_pos;
    a.pos_coord = s->pos_coord;
    a.pos_pre = s->pos_pre;
    a.pos_end = s->pos_end;
    if (!s->plans.empty()) {
        const TreePlan &tp = s->plans[s->plan];
        a.scratch = tp.scratch;
        a.huge_words = tp.huge_words;
    }
    a.seg_in_order = s->opt.segment_resolve == 1 ? 1 : 0;
    a.big_list = s->q_big_list;
    a.small_list = s->q_small_list;
    a.wg_big = s->q_wg_big;
    a.wave_small = s->q_wave_small;
    a.rc_off = s->rc_off;
    a.rc_pu = s->rc_pu;
    a.rc_pv = s->rc_pv;
    a.rc_rank = s->rc_rank;
    a.rc_coord = s->rc_coord;
    a.cand_root = s->cand_root;
    a.cand_pu = s->cand_pu;
    a.cand_pv = s->cand_pv;
    a.cand_coord = s->cand_coord;
    a.Sw = s->Sw;
    a.npad = s->npad;
    a.Tt = s->Tt;
    a.Qr = s->Qr;
    a.qbits = s->qbits;
    a.sum_cnt = s->sum_cnt;
    a.sum_rank = s->sum_rank;
    a.sum_q = s->sum_q;
    for (int L = 0; L < kMaxLevels; ++L) {
        a.lvl_slot[L] = s->lvl_slot[L];
        a.lvl_n[L] = s->lvl_n[L];
    }
    a.levels = s->levels;
    a.group = s->group;
    a.chunk_words = s->chunk_words;
    a.arrive = s->arrive;
    a.sel = s->best;
    a.blk_q = s->blk_q;
    a.blk_m = s->blk_m;
    a.blk_g = s->blk_g;
    a.C_blk = s->C_blk;
    a.D_blk = s->D_blk;
    a.Dcnt_blk = s->Dcnt_blk;
    a.ulist = s->ulist;
    a.ucount = s->ucount;
    a.Sf_blk = s->Sf_blk;
    a.C_store = s->C_store;
    a.Ct = s->Ct;
    a.Bc = s->Bc;
    a.pw_list = s->pw_list;
    a.pw_off = s->pw_off;
    a.pw_cnt = s->pw_cnt;
    a.pair_i = s->pair_i;
    a.pair_c = s->pair_c;
    a.status = s->status;
    a.upd_bytes = s->upd_bytes;
    a.K = s->K;
    a.e = e;
    a.kk = kk;
    a.N = s->N;
    a.W = s->W;
    a.n_big = s->n_qbig;
    a.n_small = s->n_qsmall;
    a.big_words = s->q_big_words;
    a.small_words = s->q_small_words;
    return a;
}

// One block of kk phases starting at e: transpose, tree pass, greedy basis, cycle vectors, update.
// An event on the stream while the update grids are timed (four per block: before and after the dot
// launch, before and after the XOR launch).
int tune_mark(mcb_solver *s) {
    if (!s->tune_ev || s->tune_n >= s->tune_cap) return MCB_OK;
    HIP_TRY(hipEventRecord(s->tune_ev[s->tune_n++], s->stream));
    return MCB_OK;
}

int launch_block(mcb_solver *s, int32_t e, int32_t kk) {
    const BlockArgs a = block_args(s, e, kk);
    const int64_t wl = (static_cast<int64_t>(e) + kk + kWave - 1) / kWave;
    timing_sample(s);
    int rc = timed_begin(s, MCB_K_TRANSPOSE);
    if (rc) return rc;
    if ((rc = launch_on(s, true, k_block_transpose, dim3(static_cast<unsigned>((wl + 3) / 4)), dim3(256), 0, a)))
        return rc;
    if ((rc = timed_end(s)) || (rc = timed_begin(s, MCB_K_PARITY_SEARCH))) return rc;
    if ((rc = launch_on(s, true, k_block_q, dim3(s->q_blocks), dim3(s->q_waves * kWave), s->q_lds, a))) return rc;
    if ((rc = timed_end(s)) || (rc = timed_begin(s, MCB_K_SELECT))) return rc;
    if ((rc = launch_on(s, s->levels <= 2, k_block_basis, dim3(static_cast<unsigned>((s->lvl_n[0] + 3) / 4)),
                        dim3(256), 0, a)))
        return rc;
    if (s->levels > 2 &&
        (rc = launch_on(s, true, k_block_cascade, dim3(static_cast<unsigned>(s->lvl_n[1])), dim3(kWave), 0, a)))
        return rc;
    if ((rc = timed_end(s)) || (rc = timed_begin(s, MCB_K_CYCLE))) return rc;
    if ((rc = launch_on(s, false, k_block_cycles_final, dim3(static_cast<unsigned>(kk + (wl + 3) / 4)), dim3(256),
                        sizeof(uint64_t) * s->W, a)))
        return rc;
    const int32_t dlim = (e + kWave - 1) / kWave;
    if ((rc = launch_on(s, true, k_block_pairs, dim3(static_cast<unsigned>(1 + (s->W + 3) / 4)), dim3(256),
                        sizeof(uint64_t) * dlim + 2 * 256 * sizeof(uint32_t), a)))
        return rc;
    if ((rc = timed_end(s)) || (rc = timed_begin(s, MCB_K_UPDATE))) return rc;
    const int64_t later = s->N - e - kk;
    if ((rc = tune_mark(s))) return rc;
    if (later > 0) {
        const size_t dlds = sizeof(uint64_t) * (kBlockK + kSegPairs) + sizeof(uint32_t) * (kSegPairs + 2 * kSegWords + 1);
        // a fixed grid: a short word list is staged once per workgroup and its waves take tiles in turn;
        // a long one is taken tile by tile by whole workgroups
        if ((rc = launch_on(s, false, k_block_dot,
                            dim3(static_cast<unsigned>(std::min<int64_t>((later + kDotCols - 1) / kDotCols,
                                                                         s->dot_wg))),
                            dim3(kDotWaves * kWave), dlds, a)))
            return rc;
    }
    if ((rc = tune_mark(s)) || (rc = tune_mark(s))) return rc;
    // (column tile, word tile) pairs over the block's own columns and the listed later ones, dealt to at
    // most xor_wg workgroups (the list's length is known only on the device)
    const int64_t most = (1 + (std::max<int64_t>(later, 0) + 255) / 256) * ((wl + kTile - 1) / kTile);
    if ((rc = launch_on(s, true, k_block_xor, dim3(static_cast<unsigned>(std::min<int64_t>(most, s->xor_wg))),
                        dim3(256), sizeof(uint64_t) * 16 * 16 * kTile, a)))
        return rc;
    if ((rc = tune_mark(s))) return rc;
    s->blk_positions += s->plans[s->plan].pos_cum.back();
    s->blk_candidates += s->plans[s->plan].cand_cum.back();
    return timed_end(s);
}

int launch_phase(mcb_solver *s, int32_t e) {
    timing_sample(s);
    int rc = timed_begin(s, MCB_K_PARITY_SEARCH);
    if (rc) return rc;
    const PhaseArgs a = phase_args(s, e);
    if ((rc = launch_on(s, true, k_phase<true, true>, dim3(s->blocks), dim3(s->wpb * kWave), s->lds_bytes, a)))
        return rc;
    return timed_end(s);
}

// Checks one problem's layout on the host: parent before child, positions, coordinates and ranks
// in range, each root's candidate ranks listed ascending.  *max_pos: the largest tree.
int validate_desc(const mcb_solver_desc *d, int32_t *max_pos) {
    if (!d || d->num_roots <= 0 || d->num_non_tree <= 0 || d->num_candidates <= 0 || d->num_positions <= 0 ||
        !d->tree_offset || !d->parent_pos || !d->pos_coord || !d->cand_root || !d->cand_pos_u || !d->cand_pos_v ||
        !d->cand_coord || !d->root_cand_offset || !d->root_cand_rank)
        return fail(MCB_E_INVALID, "invalid arguments");
    if (d->num_candidates >= static_cast<int64_t>(kNoRank)) return fail(MCB_E_INVALID, "too many candidates");
    const int32_t F = d->num_roots, N = d->num_non_tree;
    const int64_t K = d->num_candidates;
    if (d->tree_offset[0] != 0 || d->tree_offset[F] != d->num_positions) return fail(MCB_E_INVALID, "tree_offset");
    if (d->root_cand_offset[0] != 0 || d->root_cand_offset[F] != K) return fail(MCB_E_INVALID, "root_cand_offset");
    int32_t mp = 0;
    for (int32_t r = 0; r < F; ++r) {
        const int64_t a = d->tree_offset[r], b = d->tree_offset[r + 1];
        if (b <= a || b - a > 0x7FFFFFFF) return fail(MCB_E_INVALID, "empty or oversized tree");
        mp = std::max<int32_t>(mp, static_cast<int32_t>(b - a));
        if (d->parent_pos[a] != kNoParent) return fail(MCB_E_INVALID, "position 0 must be the root");
        for (int64_t i = a + 1; i < b; ++i) {
            if (d->parent_pos[i] >= static_cast<uint64_t>(i - a))
                return fail(MCB_E_INVALID, "parent position must precede its child");
            if (d->pos_coord[i] >= N) return fail(MCB_E_INVALID, "coordinate out of range");
        }
    }
    for (int64_t k = 0; k < K; ++k) {
        const uint32_t r = d->cand_root[k];
        if (r >= static_cast<uint32_t>(F)) return fail(MCB_E_INVALID, "candidate root out of range");
        const int64_t np = d->tree_offset[r + 1] - d->tree_offset[r];
        if (d->cand_pos_u[k] >= np || d->cand_pos_v[k] >= np || d->cand_coord[k] >= N)
            return fail(MCB_E_INVALID, "candidate position out of range");
    }
    for (int32_t r = 0; r < F; ++r) {
        const int64_t a = d->root_cand_offset[r], b = d->root_cand_offset[r + 1];
        if (b < a) return fail(MCB_E_INVALID, "root_cand_offset");
        for (int64_t i = a; i < b; ++i) {
            const uint32_t k = d->root_cand_rank[i];
            if (k >= K || d->cand_root[k] != static_cast<uint32_t>(r) || (i > a && k <= d->root_cand_rank[i - 1]))
                return fail(MCB_E_INVALID, "root_cand_rank must list each root's ranks ascending");
        }
    }
    *max_pos = mp;
    return MCB_OK;
}

// DFS pre-order intervals of every tree layout: pre[x] is x's index in a pre-order walk of its
// root's tree (children in position order) and end[x] = pre[x] + subtree size, so that x is an
// ancestor-or-self of y exactly when pre[x] <= pre[y] < end[x].
// Host threads for the set-up loops: OMP_NUM_THREADS when set, else the hardware's, at most 16.
int host_threads() {
    int t = static_cast<int>(std::thread::hardware_concurrency());
    if (const char *env = std::getenv("OMP_NUM_THREADS")) t = std::atoi(env);
    return std::max(1, std::min(16, t));
}

// f(i) for every i in [0, n), items dealt in chunks of `grain` to the host threads.
template <typename Fn>
void parallel_for(int64_t n, int64_t grain, Fn f) {
    const int T = host_threads();
    if (T <= 1 || n <= grain) {
        for (int64_t i = 0; i < n; ++i) f(i);
        return;
    }
    std::atomic<int64_t> next{0};
    auto work = [&]() {
        for (;;) {
            const int64_t i0 = next.fetch_add(grain);
            if (i0 >= n) return;
            const int64_t i1 = std::min(n, i0 + grain);
            for (int64_t i = i0; i < i1; ++i) f(i);
        }
    };
    std::vector<std::thread> pool;
    const int64_t want = std::min<int64_t>(T, (n + grain - 1) / grain);
    for (int64_t t = 1; t < want; ++t) pool.emplace_back(work);
    work();
    for (auto &th : pool) th.join();
}

// An allocator that leaves trivially constructible elements uninitialised on resize: the arena arrays are
// written in full by the assembly, so zero-filling gigabytes on one thread first is wasted.
template <typename T>
struct NoInit : std::allocator<T> {
    template <typename U>
    struct rebind {
        using other = NoInit<U>;
    };
    NoInit() = default;
    template <typename U>
    NoInit(const NoInit<U> &) noexcept {}
    template <typename U>
    void construct(U *p) noexcept {
        ::new (static_cast<void *>(p)) U;
    }
    template <typename U, typename... A>
    void construct(U *p, A &&...args) {
        ::new (static_cast<void *>(p)) U(std::forward<A>(args)...);
    }
};
template <typename T>
using RawVec = std::vector<T, NoInit<T>>;

// Set-up stage times on stderr when MCB_HIP_TIMING is set (seconds since the previous mark).
struct SetupClock {
    bool on = std::getenv("MCB_HIP_TIMING") != nullptr;
    std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
    void mark(const char *stage) {
        if (!on) return;
        const auto now = std::chrono::steady_clock::now();
        std::fprintf(stderr, "[mcb_hip] %-28s %8.3f s\n", stage, std::chrono::duration<double>(now - t).count());
        t = now;
    }
};

template <typename Vec>
void preorder_intervals(const int64_t *tree_offset, const uint32_t *parent_pos, int32_t F, Vec &pre, Vec &end) {
    const int64_t P = tree_offset[F];
    pre.resize(P);
    end.resize(P);
    // roots in chunks of 64, each thread with its own scratch
    parallel_for((F + 63) / 64, 1, [&](int64_t chunk) {
        thread_local std::vector<uint32_t> size, next;
        const int32_t r1 = static_cast<int32_t>(std::min<int64_t>(F, 64 * (chunk + 1)));
        for (int32_t r = static_cast<int32_t>(64 * chunk); r < r1; ++r) {
            const int64_t a = tree_offset[r];
            const uint32_t n = static_cast<uint32_t>(tree_offset[r + 1] - a);
            const uint32_t *pp = parent_pos + a;
            size.assign(n, 1);
            for (uint32_t x = n; x-- > 1;) size[pp[x]] += size[x];
            next.assign(n, 0);
            uint32_t *pr = pre.data() + a, *en = end.data() + a;
            pr[0] = 0;
            next[0] = 1;
            for (uint32_t x = 1; x < n; ++x) {  // parents precede children
                const uint32_t p = pp[x];
                pr[x] = next[p];
                next[p] += size[x];
                next[x] = pr[x] + 1;
            }
            for (uint32_t x = 0; x < n; ++x) en[x] = pr[x] + size[x];
        }
    });
}

// Lets `kernel` take `bytes` of dynamic LDS, beyond the default 64 KB when a workgroup's LDS allows
// it; MCB_E_INVALID with `what` when it cannot hold them.
int reserve_dynamic_lds(const void *kernel, int64_t bytes, const char *what) {
    hipFuncAttributes fattr{};
    HIP_TRY(hipFuncGetAttributes(&fattr, kernel));
    const int64_t lds_max =
        std::min<int64_t>(g_dev.prop.sharedMemPerBlock, 160 * 1024) - static_cast<int64_t>(fattr.sharedSizeBytes);
    if (bytes > lds_max) return fail(MCB_E_INVALID, what);
    if (bytes > kLdsLimit)
        HIP_TRY(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds_max)));
    return MCB_OK;
}

// Scratch of mcb_solver_parity_noalloc: S (W words, padded to 256 bytes), then one parity byte per tree
// position.  Each device holds its own copy until the process ends.  Every call on a device shares it and
// it keeps the last call's contents; calls that overlap on one device race.
constexpr size_t kParityScratchBytes = size_t(64) << 20;
static __device__ __align__(256) uint8_t scratch[kParityScratchBytes];

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxScratchDevices] = {};

// The device address of `scratch` on the current device, looked up at the first call on that device.
// Host code uses only this address: the array's name is a host-side stand-in, not device memory.
int scratch_address(void **out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxScratchDevices) return fail(MCB_E_DEVICE, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(MCB_E_DEVICE, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return MCB_OK;
}

// Levels below the root of every tree.
std::vector<uint32_t> tree_depths(const int64_t *tree_off, const uint32_t *parent_pos, int64_t F) {
    std::vector<uint32_t> depth(static_cast<size_t>(F), 0), level;
    for (int64_t r = 0; r < F; ++r) {
        const int64_t a0 = tree_off[r], n = tree_off[r + 1] - a0;
        level.assign(static_cast<size_t>(n), 0);
        uint32_t d = 0;
        for (int64_t x = 1; x < n; ++x) {
            level[x] = level[parent_pos[a0 + x]] + 1;
            d = std::max(d, level[x]);
        }
        depth[r] = d;
    }
    return depth;
}

// Builds the plan for nw waves per workgroup; MCB_E_INVALID when its LDS does not fit one workgroup.
// rc_off: each root's candidate run in the rc_* arrays.
// cap_words: the largest tree kept in LDS (0: as large as fits).
int plan_tree_pass(int nw, int64_t cap_words, int nsplit, const int64_t *tree_off, const int64_t *rc_off,
                   const std::vector<uint32_t> &depth, const std::vector<int64_t> &ranges, const void *kernel,
                   TreePlan &p, int64_t &bytes) {
    const size_t G = ranges.size() - 1;
    // dynamic LDS: what the kernel's own static LDS leaves of a workgroup's budget
    hipFuncAttributes fattr{};
    HIP_TRY(hipFuncGetAttributes(&fattr, kernel));
    const int64_t lds_max =
        std::min<int64_t>(g_dev.prop.sharedMemPerBlock, 160 * 1024) - static_cast<int64_t>(fattr.sharedSizeBytes);
    // largest multiple of 64 positions whose words and stamps fit the LDS, or the requested cap
    int64_t big_cap = lds_max * 64 / (8 * 64 + 4) / kWave * kWave;
    if (cap_words > 0) big_cap = std::min(big_cap, cap_words);
    int64_t largest = 0;
    for (int64_t r = 0; r < ranges.back(); ++r) largest = std::max<int64_t>(largest, tree_off[r + 1] - tree_off[r]);
    const int64_t largest_words = std::min<int64_t>((largest + kWave - 1) / kWave * kWave, big_cap);
    const int64_t small_max = std::max<int64_t>(kWave, std::min<int64_t>(1024, largest_words / nw / kWave * kWave));
    // sizes first: the LDS, hence the resident workgroups, hence the run length of split candidates
    int64_t max_big = 0, max_small = 1, max_huge = 0, total_cands = rc_off[ranges.back()] - rc_off[0];
    for (int64_t r = 0; r < ranges.back(); ++r) {
        const int64_t np = tree_off[r + 1] - tree_off[r];
        if (rc_off[r + 1] == rc_off[r]) continue;
        if (np > small_max) {
            if (np > big_cap) max_huge = std::max(max_huge, np);
            else max_big = std::max(max_big, np);
        } else {
            max_small = std::max(max_small, np);
        }
    }
    const int64_t big_words = (max_big + kWave - 1) / kWave * kWave;
    const int64_t small_words = (max_small + kWave - 1) / kWave * kWave;
    const int64_t huge_words = (max_huge + kWave - 1) / kWave * kWave;
    const int64_t lds = std::max<int64_t>({8 * big_words + 4 * (big_words / kWave), 8ll * nw * small_words,
                                           4 * (huge_words / kWave)});
    if (lds > lds_max) return MCB_E_INVALID;
    p.nw = nw;
    p.big_words = static_cast<int32_t>(big_words);
    p.small_words = static_cast<int32_t>(small_words);
    p.huge_words = huge_words;
    p.lds = static_cast<int32_t>(std::max<int64_t>(lds, 8));
    HIP_TRY(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds_max)));
    int per_cu = 0;
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, nw * kWave, p.lds));
    // a multiple of the XCD count (the batched tree pass deals workgroups to XCD sections)
    p.blocks = std::max(1, per_cu) * g_dev.prop.multiProcessorCount;
    p.blocks = (p.blocks + kXcds - 1) / kXcds * kXcds;
    const int64_t run = std::max<int64_t>(kItemCands, (total_cands + p.blocks - 1) / p.blocks);

    std::vector<std::vector<uint32_t>> big(nsplit), small(nsplit);
    std::vector<std::vector<uint8_t>> narrow(nsplit);
    std::vector<uint32_t> order, item_root;
    std::vector<int64_t> item_cb, item_ce;
    p.big_cum.assign(nsplit, std::vector<int32_t>(G + 1, 0));
    p.small_cum.assign(nsplit, std::vector<int32_t>(G + 1, 0));
    p.pos_cum.assign(G + 1, 0);
    p.cand_cum.assign(G + 1, 0);
    int64_t pos = 0, cand = 0;
    auto add_item = [&](uint32_t r, int64_t cb, int64_t ce) {
        item_root.push_back(r);
        item_cb.push_back(cb);
        item_ce.push_back(ce);
        return static_cast<uint32_t>(item_root.size() - 1);
    };
    for (size_t g = 0; g < G; ++g) {
        const int x = static_cast<int>(g % nsplit);
        order.resize(static_cast<size_t>(ranges[g + 1] - ranges[g]));
        std::iota(order.begin(), order.end(), static_cast<uint32_t>(ranges[g]));
        auto npos = [&](uint32_t r) { return tree_off[r + 1] - tree_off[r]; };
        std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return npos(a) > npos(b); });
        for (uint32_t r : order) {
            const int64_t np = npos(r), cb = rc_off[r], ce = rc_off[r + 1];
            if (ce == cb) continue;  // no candidate: nothing to emit
            cand += ce - cb;
            if (np > small_max) {
                for (int64_t c0 = cb; c0 < ce; c0 += run) {
                    big[x].push_back(add_item(r, c0, std::min(ce, c0 + run)));
                    narrow[x].push_back(np < 64ll * (depth[r] + 1) ? 1 : 0);
                    pos += np;
                }
            } else {
                small[x].push_back(add_item(r, cb, ce));
                pos += np;
            }
        }
        p.pos_cum[g + 1] = pos;
        p.cand_cum[g + 1] = cand;
        for (int y = 0; y < nsplit; ++y) {
            p.big_cum[y][g + 1] = static_cast<int32_t>(big[y].size());
            p.small_cum[y][g + 1] = static_cast<int32_t>(small[y].size());
        }
    }
    std::vector<uint32_t> big_all, small_all;
    std::vector<uint8_t> narrow_all;
    p.big_off.assign(nsplit, 0);
    p.small_off.assign(nsplit, 0);
    for (int x = 0; x < nsplit; ++x) {
        p.big_off[x] = static_cast<int32_t>(big_all.size());
        p.small_off[x] = static_cast<int32_t>(small_all.size());
        big_all.insert(big_all.end(), big[x].begin(), big[x].end());
        narrow_all.insert(narrow_all.end(), narrow[x].begin(), narrow[x].end());
        small_all.insert(small_all.end(), small[x].begin(), small[x].end());
    }
    std::vector<int32_t> wg_big, wave_small;
    if (nsplit == 1) {
        // One problem: longest-processing-time assignment under a cost model (ns of one wave, fitted
        // to per-item timestamps of this pass on Delaunay graphs): a small item 2700 + 8.8 per tree
        // position + 17 per candidate; a narrow tree walked by one wave while the others wait
        // (occupying the whole workgroup) 10.7 per position in LDS, 6.8 per position in scratch,
        // where every wave walks a segment, then its candidate run 2500 + 18.4 per candidate over
        // the nw waves; a wide (coop) tree's walk and run shared by the nw waves.  Most items are
        // small, so the per-item term dominates their cost.  Big items go, costliest first, to the
        // least loaded workgroup; then small items to the least loaded wave, each wave starting from
        // its workgroup's big-item time.
        const int nwv = nw;
        auto ncand = [&](uint32_t it) { return static_cast<double>(item_ce[it] - item_cb[it]); };
        auto npos_of = [&](uint32_t it) {
            return static_cast<double>(tree_off[item_root[it] + 1] - tree_off[item_root[it]]);
        };
        std::vector<double> bcost(big_all.size());
        for (size_t i = 0; i < big_all.size(); ++i) {
            const uint32_t it = big_all[i];
            const double np = npos_of(it);
            bcost[i] = narrow_all[i] ? (np > big_words ? 6.8 : 10.7) * np + 2500.0 + 18.4 * ncand(it) / nwv
                                     : 2700.0 + (8.8 * np + 17.0 * ncand(it)) / nwv;
        }
        std::vector<size_t> bo(big_all.size());
        std::iota(bo.begin(), bo.end(), size_t{0});
        std::stable_sort(bo.begin(), bo.end(), [&](size_t x, size_t y) { return bcost[x] > bcost[y]; });
        using Load = std::pair<double, int32_t>;
        std::priority_queue<Load, std::vector<Load>, std::greater<Load>> heap;
        for (int32_t w = 0; w < p.blocks; ++w) heap.push({0.0, w});
        std::vector<std::vector<size_t>> per_wg(static_cast<size_t>(p.blocks));
        std::vector<double> wg_load(static_cast<size_t>(p.blocks), 0.0);
        for (size_t i : bo) {
            Load l = heap.top();
            heap.pop();
            per_wg[static_cast<size_t>(l.second)].push_back(i);
            l.first += bcost[i];
            wg_load[static_cast<size_t>(l.second)] = l.first;
            heap.push(l);
        }
        std::vector<uint32_t> nb;
        std::vector<uint8_t> nn;
        wg_big.push_back(0);
        for (const auto &v : per_wg) {
            for (size_t i : v) {
                nb.push_back(big_all[i]);
                nn.push_back(narrow_all[i]);
            }
            wg_big.push_back(static_cast<int32_t>(nb.size()));
        }
        big_all.swap(nb);
        narrow_all.swap(nn);
        std::vector<double> scost(small_all.size());
        for (size_t i = 0; i < small_all.size(); ++i)
            scost[i] = 2700.0 + 8.8 * npos_of(small_all[i]) + 17.0 * ncand(small_all[i]);
        std::vector<size_t> so(small_all.size());
        std::iota(so.begin(), so.end(), size_t{0});
        std::stable_sort(so.begin(), so.end(), [&](size_t x, size_t y) { return scost[x] > scost[y]; });
        const size_t waves = static_cast<size_t>(p.blocks) * static_cast<size_t>(nwv);
        std::priority_queue<Load, std::vector<Load>, std::greater<Load>> wheap;
        for (size_t c = 0; c < waves; ++c) wheap.push({wg_load[c / nwv], static_cast<int32_t>(c)});
        std::vector<std::vector<size_t>> per_wave(waves);
        for (size_t i : so) {
            Load l = wheap.top();
            wheap.pop();
            per_wave[static_cast<size_t>(l.second)].push_back(i);
            l.first += scost[i];
            wheap.push(l);
        }
        std::vector<uint32_t> ns;
        wave_small.push_back(0);
        for (const auto &v : per_wave) {
            for (size_t i : v) ns.push_back(small_all[i]);
            wave_small.push_back(static_cast<int32_t>(ns.size()));
        }
        small_all.swap(ns);
    }
    p.n_big = static_cast<int32_t>(big_all.size());
    p.n_small = static_cast<int32_t>(small_all.size());
    auto meta_of = [&](uint32_t it, bool nar) {
        const uint32_t r = item_root[it];
        ItemMeta m{};
        m.off = tree_off[r];
        m.cb = item_cb[it];
        m.ce = item_ce[it];
        m.npos = static_cast<uint32_t>(tree_off[r + 1] - tree_off[r]) | (nar ? kNarrowBit : 0u);
        m.root = r;
        return m;
    };
    std::vector<ItemMeta> big_meta(big_all.size()), small_meta(small_all.size());
    for (size_t i = 0; i < big_all.size(); ++i) big_meta[i] = meta_of(big_all[i], narrow_all[i] != 0);
    for (size_t i = 0; i < small_all.size(); ++i) small_meta[i] = meta_of(small_all[i], false);
    int rc;
    if ((rc = upload(&p.big, big_meta.data(), big_meta.size(), bytes)) ||
        (rc = upload(&p.small, small_meta.data(), small_meta.size(), bytes)))
        return rc;
    if (nsplit == 1 && ((rc = upload(&p.wg_big, wg_big.data(), wg_big.size(), bytes)) ||
                        (rc = upload(&p.wave_small, wave_small.data(), wave_small.size(), bytes))))
        return rc;
    return MCB_OK;
}

// The plans worth trying: the requested width, or 16, 8 and 4 waves, each with the largest LDS-resident
// tree as large as fits, 8192 or 2048 positions (trees above the cap keep their words in scratch: less LDS
// per workgroup, more workgroups per CU); caps that change nothing are left out.  *scratch receives one
// buffer large enough for every plan (one runs at a time), owned by the caller.
int make_plans(int want, int want_cap, int nsplit, const int64_t *tree_off, const uint32_t *parent_pos,
               const int64_t *rc_off, const std::vector<int64_t> &ranges, const void *kernel,
               std::vector<TreePlan> &plans, int64_t &bytes, uint64_t **scratch) {
    if (want != 0 && want != 4 && want != 8 && want != 16)
        return fail(MCB_E_INVALID, "waves_per_block must be 0, 4, 8 or 16");
    if (want_cap < -1 || (want_cap > 0 && want_cap % kWave))
        return fail(MCB_E_INVALID, "tree_cap must be -1, 0 or a positive multiple of 64");
    const std::vector<uint32_t> depth = tree_depths(tree_off, parent_pos, ranges.back());
    int64_t largest = 0;
    for (int64_t r = 0; r < ranges.back(); ++r) largest = std::max<int64_t>(largest, tree_off[r + 1] - tree_off[r]);
    for (int nw : {16, 8, 4}) {
        if (want && nw != want) continue;
        for (int64_t cap : {int64_t{0}, int64_t{8192}, int64_t{2048}}) {
            if (want_cap == -1 && cap) continue;
            if (want_cap > 0) {
                if (cap) continue;
                cap = want_cap;
            } else if (cap && cap >= largest) {
                continue;  // the same plan as without a cap
            }
            TreePlan p;
            const int rc = plan_tree_pass(nw, cap, nsplit, tree_off, rc_off, depth, ranges, kernel, p, bytes);
            if (rc == MCB_OK) plans.push_back(std::move(p));
            else {
                free_plan(p);
                if (rc != MCB_E_INVALID) return rc;
            }
        }
    }
    int64_t words = 0;
    // every workgroup's path words, then every workgroup's anchors (u32) for the segmented walk
    for (const TreePlan &p : plans)
        words = std::max<int64_t>(words, static_cast<int64_t>(p.blocks) * p.huge_words * 3 / 2 + 1);
    *scratch = nullptr;
    if (words) {
        int rc = dev_alloc(scratch, static_cast<size_t>(words), bytes);
        if (rc) return rc;
    }
    for (TreePlan &p : plans) p.scratch = p.huge_words ? *scratch : nullptr;
    return MCB_OK;
}

// Makes plan i the tree pass's launch shape.
void use_plan(mcb_solver *s, size_t i) {
    const TreePlan &p = s->plans[i];
    s->plan = static_cast<int32_t>(i);
    s->q_waves = p.nw;
    s->q_big_words = p.big_words;
    s->q_small_words = p.small_words;
    s->q_lds = p.lds;
    s->q_blocks = p.blocks;
    s->n_qbig = p.n_big;
    s->n_qsmall = p.n_small;
    s->q_big_list = p.big;
    s->q_small_list = p.small;
    s->q_wg_big = p.wg_big;
    s->q_wave_small = p.wave_small;
}


// Times `launch` (one tree pass under the plan `use(i)` selects) for every plan and leaves the fastest
// selected: the first of three runs warms up, the best of the other two counts.
template <typename UseFn, typename LaunchFn>
int pick_fastest(size_t nplans, hipStream_t stream, UseFn use, LaunchFn launch) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        if (e0) (void)hipEventDestroy(e0);
        return fail(MCB_E_DEVICE, "hipEventCreate");
    }
    float best = 0.0f;
    size_t bi = 0;
    int rc = MCB_OK;
    for (size_t i = 0; i < nplans && rc == MCB_OK; ++i) {
        use(i);
        float t = 0.0f;
        for (int rep = 0; rep < 3; ++rep) {
            float ms = 0.0f;
            if (hipEventRecord(e0, stream) != hipSuccess || (rc = launch()) != MCB_OK ||
                hipEventRecord(e1, stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                if (rc == MCB_OK) rc = fail(MCB_E_DEVICE, "tree pass timing");
                break;
            }
            if (rep == 1 || (rep == 2 && ms < t)) t = ms;
        }
        if (rc == MCB_OK && (i == 0 || t < best)) {
            best = t;
            bi = i;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    use(bi);
    return MCB_OK;
}

// Grids of the support vector update: each of dot_workgroups and xor_workgroups that the options leave at
// 0 is timed over the first (at most eight) blocks of phases, from the unit support vectors, among three
// grid caps, and the fastest kept; the other keeps its option's value.  A run with the middle caps warms up
// first and counts for nothing.  Problems of fewer than 256 blocks keep the middle caps: there the timing
// would cost more than a loop could gain from it.  Every grid gives the same results (both kernels stride
// over their tiles); only the time differs.  Leaves the solver reset.
int tune_update_grids(mcb_solver *s) {
    static const int32_t kDotTry[3] = {512, static_cast<int32_t>(kDotWorkgroups), 2048};
    static const int32_t kXorTry[3] = {1024, static_cast<int32_t>(kXorWorkgroups), 4096};
    const bool td = s->opt.dot_workgroups == 0, tx = s->opt.xor_workgroups == 0;
    s->dot_wg = td ? kDotTry[1] : s->opt.dot_workgroups;
    s->xor_wg = tx ? kXorTry[1] : s->opt.xor_workgroups;
    const int32_t nb = s->bk > 1 ? std::min<int32_t>(8, s->N / s->bk) : 0;
    if ((!td && !tx) || s->bk <= 1 || s->N / s->bk < 256) return MCB_OK;
    std::vector<hipEvent_t> ev(4 * static_cast<size_t>(nb), nullptr);
    int rc = MCB_OK;
    for (auto &x : ev)
        if (hipEventCreate(&x) != hipSuccess) {
            rc = fail(MCB_E_DEVICE, "hipEventCreate");
            break;
        }
    const int32_t timing = s->opt.timing;
    s->opt.timing = 0;
    double dt[3] = {0, 0, 0}, xt[3] = {0, 0, 0};
    const int order[4] = {1, 0, 1, 2};
    for (int r = 0; r < 4 && rc == MCB_OK; ++r) {
        const int c = order[r];
        if (td) s->dot_wg = kDotTry[c];
        if (tx) s->xor_wg = kXorTry[c];
        if ((rc = mcb_solver_reset(s))) break;
        s->tune_ev = ev.data();
        s->tune_n = 0;
        s->tune_cap = ev.size();
        rc = mcb_solver_run(s, nb * s->bk);
        s->tune_ev = nullptr;
        if (rc) break;
        if (hipStreamSynchronize(s->stream) != hipSuccess || s->tune_n != ev.size()) {
            rc = fail(MCB_E_DEVICE, "update grid timing");
            break;
        }
        if (r == 0) continue;
        for (int32_t b = 0; b < nb; ++b) {
            float d = 0.0f, x = 0.0f;
            if (hipEventElapsedTime(&d, ev[4 * b], ev[4 * b + 1]) != hipSuccess ||
                hipEventElapsedTime(&x, ev[4 * b + 2], ev[4 * b + 3]) != hipSuccess) {
                rc = fail(MCB_E_DEVICE, "update grid timing");
                break;
            }
            dt[c] += d;
            xt[c] += x;
        }
    }
    s->opt.timing = timing;
    for (auto &x : ev)
        if (x) (void)hipEventDestroy(x);
    if (rc) return rc;
    int bd = 0, bx = 0;
    for (int c = 1; c < 3; ++c) {
        if (dt[c] < dt[bd]) bd = c;
        if (xt[c] < xt[bx]) bx = c;
    }
    if (td) s->dot_wg = kDotTry[bd];
    if (tx) s->xor_wg = kXorTry[bx];
    return mcb_solver_reset(s);
}

// Buffers and launch shape of the blocked form.  Chooses the per-phase form instead when
// block_phases asks for it or (block_phases 0) when no tree pass plan fits the LDS.
int setup_blocked(mcb_solver *s, const int64_t *tree_offset, const uint32_t *parent_pos, const int64_t *rc_off) {
    const int want = s->opt.block_phases;
    if (want < 0 || want > kBlockK) return fail(MCB_E_INVALID, "block_phases must be 0..64");
    if (want == 1) return MCB_OK;
    const std::vector<int64_t> ranges{0, static_cast<int64_t>(s->F)};
    int rc = make_plans(s->opt.waves_per_block, s->opt.tree_cap, 1, tree_offset, parent_pos, rc_off, ranges,
                        reinterpret_cast<const void *>(k_block_q), s->plans, s->bytes, &s->q_scratch);
    if (rc) return rc;
    if (s->plans.empty()) {
        if (want > 1 || s->opt.waves_per_block)
            return fail(MCB_E_INVALID, "a tree is too large for the blocked form's LDS budget");
        return MCB_OK;  // automatic choice: fall back to one kernel per phase
    }
    // the block's cycle vectors (8 B per word of C) and its pair list (8 B per word below e plus two
    // 256-entry scans) stage in LDS: up to about 1.29 M coordinates in a 160 KB workgroup
    if ((rc = reserve_dynamic_lds(reinterpret_cast<const void *>(k_block_cycles_final), 8ll * s->W,
                                  "support vector too long for the cycle kernel's LDS")) ||
        (rc = reserve_dynamic_lds(reinterpret_cast<const void *>(k_block_pairs), 8ll * s->W + 2048,
                                  "support vector too long for the pair kernel's LDS")))
        return rc;
    s->bk = want > 1 ? want : kBlockK;
    use_plan(s, 0);
    // greedy-basis levels: chunk_words bitmap words per first-level basis, group bases per merge.  When
    // at most kMaxRounds rounds of 64 bitmap words per first-level basis bring the first level within
    // one merge group of 64 (Delaunay 100k: 11 rounds), a first-level basis reads that many: the last
    // one to finish then merges them all and finalises the block in the same launch.  Otherwise (and
    // with a chosen merge group) one round, 4096 ranks, per first-level basis: a first-level wave
    // reading many rounds of a dense bitmap (ER 20 000) inserts batch after batch alone, which costs
    // more than the merge levels it saves.
    int64_t n = (s->K + kChunk - 1) / kChunk, slots = 0;
    s->chunk_words = kWave;
    if (s->opt.merge_group == 0 && n > 64 && n <= 64 * kMaxRounds) {
        const int64_t rounds = (n + 63) / 64;
        s->chunk_words = static_cast<int32_t>(rounds * kWave);
        n = (s->K + static_cast<int64_t>(kChunk) * rounds - 1) / (static_cast<int64_t>(kChunk) * rounds);
    }
    s->group = merge_group_size(n, s->opt.merge_group);
    s->levels = 0;
    for (;;) {
        if (s->levels >= kMaxLevels) return fail(MCB_E_INVALID, "too many candidates for the basis levels");
        s->lvl_slot[s->levels] = slots;
        s->lvl_n[s->levels] = n;
        slots += n;
        ++s->levels;
        if (n == 1) break;
        n = (n + s->group - 1) / s->group;
    }
    int64_t &B = s->bytes;
    if ((rc = dev_alloc(&s->Tt, static_cast<size_t>(s->npad), B))) return rc;
    if ((rc = dev_alloc(&s->Qr, static_cast<size_t>(s->K), B))) return rc;
    if ((rc = dev_alloc(&s->qbits, static_cast<size_t>((s->K + 63) / 64), B))) return rc;
    if ((rc = dev_alloc(&s->sum_cnt, static_cast<size_t>(slots), B))) return rc;
    if ((rc = dev_alloc(&s->arrive, static_cast<size_t>(slots), B))) return rc;
    HIP_TRY(hipMemset(s->arrive, 0, static_cast<size_t>(slots) * sizeof(uint32_t)));
    if ((rc = dev_alloc(&s->sum_rank, static_cast<size_t>(slots) * kWave, B))) return rc;
    if ((rc = dev_alloc(&s->sum_q, static_cast<size_t>(slots) * kWave, B))) return rc;
    if ((rc = dev_alloc(&s->blk_q, kBlockK, B))) return rc;
    if ((rc = dev_alloc(&s->blk_m, kBlockK, B))) return rc;
    if ((rc = dev_alloc(&s->blk_g, kBlockK, B))) return rc;
    if ((rc = dev_alloc(&s->C_blk, static_cast<size_t>(kBlockK) * s->W, B))) return rc;
    if ((rc = dev_alloc(&s->Sf_blk, static_cast<size_t>(kBlockK) * s->W, B))) return rc;
    if ((rc = dev_alloc(&s->D_blk, static_cast<size_t>(kBlockK) * s->W, B))) return rc;
    if ((rc = dev_alloc(&s->Dcnt_blk, kBlockK, B))) return rc;
    if ((rc = dev_alloc(&s->ulist, static_cast<size_t>(s->npad), B))) return rc;
    if ((rc = dev_alloc(&s->ucount, 1, B))) return rc;
    HIP_TRY(hipMemset(s->ucount, 0, sizeof(uint32_t)));
    if ((rc = dev_alloc(&s->upd_bytes, kByteSlots, B))) return rc;
    if ((rc = dev_alloc(&s->Ct, static_cast<size_t>(s->npad), B))) return rc;
    if ((rc = dev_alloc(&s->Bc, static_cast<size_t>(s->npad), B))) return rc;
    if ((rc = dev_alloc(&s->pw_list, static_cast<size_t>(s->W), B))) return rc;
    if ((rc = dev_alloc(&s->pw_off, static_cast<size_t>(s->W) + 1, B))) return rc;
    if ((rc = dev_alloc(&s->pw_cnt, 1, B))) return rc;
    if ((rc = dev_alloc(&s->pair_i, static_cast<size_t>(kBlockK) * s->W, B))) return rc;
    if ((rc = dev_alloc(&s->pair_c, static_cast<size_t>(kBlockK) * s->W, B))) return rc;
    return MCB_OK;
}

}  // namespace

extern "C" {

void mcb_solver_options_init(mcb_solver_options *opt) {
    if (!opt) return;
    std::memset(opt, 0, sizeof(*opt));
}

void mcb_solver_destroy(mcb_solver *s) {
    if (!s) return;
    void *ptrs[] = {s->tree_off, s->parent_pos, s->pos_pre, s->pos_end, s->pos_coord, s->big_list, s->small_list,
                    s->root_min_rank, s->rc_off, s->rc_pu, s->rc_pv, s->rc_rank, s->rc_coord, s->cand_root,
                    s->cand_pu, s->cand_pv, s->cand_coord, s->Sw, s->Scur, s->C, s->C_store, s->D, s->Dcount,
                    s->best, s->done, s->status, s->wave_stats, s->flag, s->E, s->Ecount, s->cdot,
                    s->Tt, s->Qr, s->qbits, s->sum_q, s->blk_q, s->blk_m, s->blk_g, s->C_blk,
                    s->Sf_blk, s->sum_cnt, s->arrive, s->sum_rank, s->D_blk, s->Dcnt_blk, s->upd_bytes,
                    s->Ct, s->Bc, s->pair_c, s->pw_list, s->pw_off, s->pw_cnt, s->pair_i, s->ulist, s->ucount};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (TreePlan &p : s->plans) free_plan(p);
    if (s->q_scratch) (void)hipFree(s->q_scratch);
    if (s->graph_exec) (void)hipGraphExecDestroy(s->graph_exec);
    if (s->graph) (void)hipGraphDestroy(s->graph);
    for (hipEvent_t e : s->ev_pool) (void)hipEventDestroy(e);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

int mcb_solver_create(const mcb_solver_desc *d, const mcb_solver_options *opt, mcb_solver **out) {
    if (!out) return fail(MCB_E_INVALID, "invalid arguments");
    *out = nullptr;
    int32_t max_pos = 0;
    int rc = validate_desc(d, &max_pos);
    if (rc) return rc;
    if (opt && (opt->merge_group < 0 || opt->merge_group == 1 || opt->merge_group > 64))
        return fail(MCB_E_INVALID, "merge_group must be 0 or 2..64");
    if (opt && opt->timing_every < 0) return fail(MCB_E_INVALID, "timing_every must be 0 or more");
    if (opt && (opt->segment_resolve < 0 || opt->segment_resolve > 1))
        return fail(MCB_E_INVALID, "segment_resolve must be 0 or 1");
    if (opt && (opt->dot_workgroups < 0 || opt->dot_workgroups > 65536 || opt->xor_workgroups < 0 ||
                opt->xor_workgroups > 65536))
        return fail(MCB_E_INVALID, "dot_workgroups and xor_workgroups must be 0..65536");
    rc = ensure_device();
    if (rc) return rc;
    mcb_solver *s = new (std::nothrow) mcb_solver();
    if (!s) return fail(MCB_E_NOMEM, "out of memory");
    auto cleanup = [&](int code) {
        mcb_solver_destroy(s);
        return code;
    };
    if (opt) s->opt = *opt;
    s->F = d->num_roots;
    s->N = d->num_non_tree;
    s->K = d->num_candidates;
    s->Ptot = d->num_positions;
    s->W = (s->N + 63) / 64;
    s->npad = (static_cast<int64_t>(s->N) + 63) & ~static_cast<int64_t>(63);
    const int32_t F = s->F;
    const int64_t K = s->K;
    s->max_pos = max_pos;
    // Root-major candidate copies, lowest rank per root, and the two work lists (largest first).
    std::vector<uint32_t> rc_pu(K), rc_pv(K), rc_rank(K), min_rank(F, kNoRank), order(F);
    std::vector<int32_t> rc_coord(K);
    for (int32_t r = 0; r < F; ++r) {
        const int64_t a = d->root_cand_offset[r], b = d->root_cand_offset[r + 1];
        for (int64_t i = a; i < b; ++i) {
            const uint32_t k = d->root_cand_rank[i];
            rc_pu[i] = d->cand_pos_u[k];
            rc_pv[i] = d->cand_pos_v[k];
            rc_rank[i] = k;
            rc_coord[i] = d->cand_coord[k];
        }
        if (b > a) min_rank[r] = d->root_cand_rank[a];
    }
    std::iota(order.begin(), order.end(), 0u);
    auto npos_of = [&](uint32_t r) { return d->tree_offset[r + 1] - d->tree_offset[r]; };
    std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return npos_of(x) > npos_of(y); });
    std::vector<uint32_t> big, small;
    int64_t max_small = 1;
    for (uint32_t r : order) {
        if (npos_of(r) > static_cast<int64_t>(kBigPositions)) {
            big.push_back(r);
        } else {
            small.push_back(r);
            max_small = std::max(max_small, npos_of(r));
        }
    }
    s->n_big = static_cast<int32_t>(big.size());
    s->n_small = static_cast<int32_t>(small.size());
    s->p_words_big = big.empty() ? 0 : static_cast<int32_t>((s->max_pos + 63) / 64);
    s->p_words_small = small.empty() ? 0 : static_cast<int32_t>((max_small + 63) / 64);
    std::vector<uint32_t> pre, end;
    preorder_intervals(d->tree_offset, d->parent_pos, F, pre, end);

    int64_t &B = s->bytes;
    if ((rc = upload(&s->tree_off, d->tree_offset, F + 1, B))) return cleanup(rc);
    if ((rc = upload(&s->parent_pos, d->parent_pos, s->Ptot, B))) return cleanup(rc);
    if ((rc = upload(&s->pos_coord, d->pos_coord, s->Ptot, B))) return cleanup(rc);
    if ((rc = upload(&s->pos_pre, pre.data(), s->Ptot, B))) return cleanup(rc);
    if ((rc = upload(&s->pos_end, end.data(), s->Ptot, B))) return cleanup(rc);
    if ((rc = upload(&s->big_list, big.data(), big.size(), B))) return cleanup(rc);
    if ((rc = upload(&s->small_list, small.data(), small.size(), B))) return cleanup(rc);
    if ((rc = upload(&s->root_min_rank, min_rank.data(), F, B))) return cleanup(rc);
    if ((rc = upload(&s->rc_off, d->root_cand_offset, F + 1, B))) return cleanup(rc);
    if ((rc = upload(&s->rc_pu, rc_pu.data(), K, B))) return cleanup(rc);
    if ((rc = upload(&s->rc_pv, rc_pv.data(), K, B))) return cleanup(rc);
    if ((rc = upload(&s->rc_rank, rc_rank.data(), K, B))) return cleanup(rc);
    if ((rc = upload(&s->rc_coord, rc_coord.data(), K, B))) return cleanup(rc);
    if ((rc = upload(&s->cand_root, d->cand_root, K, B))) return cleanup(rc);
    if ((rc = upload(&s->cand_pu, d->cand_pos_u, K, B))) return cleanup(rc);
    if ((rc = upload(&s->cand_pv, d->cand_pos_v, K, B))) return cleanup(rc);
    if ((rc = upload(&s->cand_coord, d->cand_coord, K, B))) return cleanup(rc);
    if ((rc = dev_alloc(&s->Sw, static_cast<size_t>(s->W) * s->npad, B))) return cleanup(rc);
    if ((rc = dev_alloc(&s->Scur, 4 * static_cast<size_t>(s->W), B))) return cleanup(rc);
    if ((rc = dev_alloc(&s->C, 2 * static_cast<size_t>(s->W), B))) return cleanup(rc);
    if (s->opt.store_cycles && (rc = dev_alloc(&s->C_store, static_cast<size_t>(s->N) * s->W, B))) return cleanup(rc);
    if ((rc = dev_alloc(&s->D, 2 * static_cast<size_t>(s->W), B))) return cleanup(rc);
    if ((rc = dev_alloc(&s->Dcount, 2, B))) return cleanup(rc);
    if ((rc = dev_alloc(&s->best, s->N, B))) return cleanup(rc);
    if ((rc = dev_alloc(&s->done, s->N, B))) return cleanup(rc);
    if ((rc = dev_alloc(&s->flag, s->npad, B))) return cleanup(rc);
    if ((rc = dev_alloc(&s->E, 4 * static_cast<size_t>(s->W), B))) return cleanup(rc);
    if ((rc = dev_alloc(&s->Ecount, 4, B))) return cleanup(rc);
    if ((rc = dev_alloc(&s->cdot, 2, B))) return cleanup(rc);
    if ((rc = dev_alloc(&s->status, 1, B))) return cleanup(rc);
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
        return cleanup(fail(MCB_E_DEVICE, "hipStreamCreate"));

    // LDS of the phase kernel: S_e, one bit per position of the block's big tree plus its ready
    // stamps, and one bit per position of each wave's small tree; the epilogue reuses it for C_e and
    // its word list (12 bytes per word of C_e).
    const int64_t tree_bytes = 8ll * s->p_words_big + 8ll * ((s->p_words_big + 1) / 2) +
                               8ll * kBlockWaves * s->p_words_small;
    // Only the per-phase form (and mcb_solver_parity) needs it: 12 W bytes of LDS cap that form at
    // about 349 000 coordinates, while the blocked form takes up to 8 W + 2 KB of a 160 KB workgroup.
    const int64_t need = std::max<int64_t>(8ll * s->W + tree_bytes, 12ll * s->W);
    const bool phase_fits = need <= kLdsLimit;
    if (!phase_fits && s->opt.block_phases == 1)
        return cleanup(fail(MCB_E_INVALID, "support vector and tree layout exceed the per-phase kernel's LDS budget"));
    s->lds_bytes = phase_fits ? static_cast<int>(need) : 0;
    int64_t blocks = 1;
    if (phase_fits) {
        int per_cu = 0;
        HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_phase<true, true>, kBlockWaves * kWave,
                                                             s->lds_bytes));
        per_cu = std::max(1, per_cu);
        blocks = static_cast<int64_t>(per_cu) * g_dev.prop.multiProcessorCount;
        if (s->opt.max_blocks > 0) blocks = std::min<int64_t>(blocks, s->opt.max_blocks);
        blocks = std::min<int64_t>(blocks, s->n_big + (s->n_small + kBlockWaves - 1) / kBlockWaves);
    }
    s->blocks = static_cast<int>(std::max<int64_t>(1, blocks));
    if ((rc = dev_alloc(&s->wave_stats, 3 * static_cast<size_t>(s->blocks) * s->wpb, B))) return cleanup(rc);
    if ((rc = setup_blocked(s, d->tree_offset, d->parent_pos, d->root_cand_offset))) return cleanup(rc);
    if (s->bk <= 1 && !phase_fits)
        return cleanup(fail(MCB_E_INVALID, "a tree exceeds the blocked form's LDS budget and the support vector "
                                           "the per-phase kernel's"));
    if (s->opt.timing) {
        s->ev_pool.resize(4096);
        s->ev_kind.resize(2048);
        for (auto &e : s->ev_pool)
            if (hipEventCreate(&e) != hipSuccess) return cleanup(fail(MCB_E_DEVICE, "hipEventCreate"));
    }
    rc = mcb_solver_reset(s);
    if (rc) return cleanup(rc);
    if (s->plans.size() > 1) {
        // tree pass width: time every plan on the first block (from the unit support vectors)
        const int32_t kk = std::min(s->bk, s->N);
        const BlockArgs a = block_args(s, 0, kk);
        const unsigned words = static_cast<unsigned>((kk + kWave - 1) / kWave);
        hipLaunchKernelGGL(k_block_transpose, dim3((words + 3) / 4), dim3(256), 0, s->stream, a);
        if (hipGetLastError() != hipSuccess) return cleanup(fail(MCB_E_DEVICE, "k_block_transpose"));
        rc = pick_fastest(
            s->plans.size(), s->stream, [&](size_t i) { use_plan(s, i); },
            [&]() -> int {
                const BlockArgs b = block_args(s, 0, kk);
                hipLaunchKernelGGL(k_block_q, dim3(s->q_blocks), dim3(s->q_waves * kWave), s->q_lds, s->stream, b);
                HIP_TRY(hipGetLastError());
                return MCB_OK;
            });
        if (rc || (rc = mcb_solver_reset(s))) return cleanup(rc);
    }
    if ((rc = tune_update_grids(s))) return cleanup(rc);
    *out = s;
    return MCB_OK;
}

int mcb_solver_reset(mcb_solver *s) {
    if (!s) return fail(MCB_E_INVALID, "invalid arguments");
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipMemsetAsync(s->Sw, 0, sizeof(uint64_t) * static_cast<size_t>(s->W) * s->npad, s->stream));
    std::vector<uint64_t> s0(4 * static_cast<size_t>(s->W), 0);
    s0[0] = 1;  // S_0 = e_0 in slot 0 of the ring
    HIP_TRY(hipMemcpyAsync(s->Scur, s0.data(), sizeof(uint64_t) * s0.size(), hipMemcpyHostToDevice, s->stream));
    hipLaunchKernelGGL(k_fill_u32, dim3(static_cast<unsigned>((s->N + 255) / 256)), dim3(256), 0, s->stream, s->best,
                       kNoRank, static_cast<int64_t>(s->N));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemsetAsync(s->done, 0, sizeof(uint32_t) * static_cast<size_t>(s->N), s->stream));
    HIP_TRY(hipMemsetAsync(s->flag, 0, static_cast<size_t>(s->npad), s->stream));
    HIP_TRY(hipMemsetAsync(s->cdot, 0, 2 * sizeof(uint32_t), s->stream));
    // S_0 = e_0: one nonzero word, word 0, in slot 0 of the ring
    HIP_TRY(hipMemsetAsync(s->E, 0, 4 * sizeof(uint32_t) * s->W, s->stream));
    const uint32_t ecount0[4] = {1, 0, 0, 0};
    HIP_TRY(hipMemcpyAsync(s->Ecount, ecount0, sizeof(ecount0), hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemsetAsync(s->Dcount, 0, 2 * sizeof(uint32_t), s->stream));
    HIP_TRY(hipMemsetAsync(s->C, 0, 2 * sizeof(uint64_t) * s->W, s->stream));
    HIP_TRY(hipMemsetAsync(s->status, 0, sizeof(uint32_t), s->stream));
    HIP_TRY(hipMemsetAsync(s->wave_stats, 0, sizeof(uint64_t) * 3 * s->blocks * s->wpb, s->stream));
    if (s->bk > 1) {
        HIP_TRY(hipMemsetAsync(s->Tt, 0, sizeof(uint64_t) * static_cast<size_t>(s->npad), s->stream));
        HIP_TRY(hipMemsetAsync(s->qbits, 0, sizeof(uint64_t) * static_cast<size_t>((s->K + 63) / 64), s->stream));
        HIP_TRY(hipMemsetAsync(s->upd_bytes, 0, sizeof(unsigned long long) * kByteSlots, s->stream));
    }
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->phase = 0;
    s->blk_positions = 0;
    s->blk_candidates = 0;
    s->ev_used = 0;
    s->timing_seq = 0;
    for (int i = 0; i < MCB_K_COUNT; ++i) {
        s->kt_ms[i] = 0;
        s->kt_n[i] = 0;
    }
    return MCB_OK;
}

int mcb_solver_run(mcb_solver *s, int32_t phases) {
    if (!s || phases < 0) return fail(MCB_E_INVALID, "invalid arguments");
    const int32_t end = std::min(s->N, s->phase + phases);
    while (s->phase < end) {
        const int32_t e = s->phase;
        const int32_t kk = std::min(s->bk, end - e);
        int rc = s->bk > 1 ? launch_block(s, e, kk) : launch_phase(s, e);
        if (rc) return rc;
        s->phase = e + kk;
    }
    return MCB_OK;
}

int mcb_solver_run_graph(mcb_solver *s) {
    if (!s) return fail(MCB_E_INVALID, "invalid arguments");
    if (s->phase != 0) return fail(MCB_E_STATE, "the graph runs the whole loop from phase 0 (reset first)");
    if (s->opt.timing) return fail(MCB_E_STATE, "graph replay needs timing off");
    if (!s->graph_exec) {
        HIP_TRY(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
        int rc = MCB_OK;
        for (int32_t e = 0; e < s->N && rc == MCB_OK;) {
            const int32_t kk = std::min(s->bk, s->N - e);
            rc = s->bk > 1 ? launch_block(s, e, kk) : launch_phase(s, e);
            e += kk;
        }
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(s->stream, &g);
        if (rc || ec != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return rc ? rc : fail(MCB_E_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
        }
        s->graph = g;
        HIP_TRY(hipGraphInstantiate(&s->graph_exec, g, nullptr, nullptr, 0));
        // the counters were zero at phase 0; capture ran the host side of every launch once
        s->graph_positions = s->blk_positions;
        s->graph_candidates = s->blk_candidates;
    }
    HIP_TRY(hipGraphLaunch(s->graph_exec, s->stream));
    s->blk_positions = s->graph_positions;
    s->blk_candidates = s->graph_candidates;
    s->phase = s->N;
    return MCB_OK;
}

int mcb_solver_sync(mcb_solver *s) {
    if (!s) return fail(MCB_E_INVALID, "invalid arguments");
    HIP_TRY(hipStreamSynchronize(s->stream));
    int rc = harvest_events(s);
    if (rc) return rc;
    uint32_t st = 0;
    HIP_TRY(hipMemcpy(&st, s->status, sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (st) return fail(MCB_E_NOT_FOUND, "a phase found no odd candidate");
    return MCB_OK;
}

int32_t mcb_solver_phase(const mcb_solver *s) { return s ? s->phase : -1; }

int mcb_solver_selected(mcb_solver *s, int32_t *out, int32_t count) {
    if (!s || !out || count < 0 || count > s->N) return fail(MCB_E_INVALID, "invalid arguments");
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipMemcpy(out, s->best, sizeof(uint32_t) * count, hipMemcpyDeviceToHost));
    return MCB_OK;
}

int mcb_solver_support_vectors(mcb_solver *s, uint64_t *out) {
    if (!s || !out) return fail(MCB_E_INVALID, "invalid arguments");
    HIP_TRY(hipStreamSynchronize(s->stream));
    std::vector<uint64_t> sw(static_cast<size_t>(s->W) * s->npad);
    HIP_TRY(hipMemcpy(sw.data(), s->Sw, sizeof(uint64_t) * sw.size(), hipMemcpyDeviceToHost));
    for (int32_t j = 0; j < s->N; ++j) {
        uint64_t *row = out + static_cast<size_t>(j) * s->W;
        for (int32_t w = 0; w < s->W; ++w) row[w] = sw[static_cast<size_t>(w) * s->npad + j];
        row[j >> 6] |= 1ull << (j & 63);
    }
    return MCB_OK;
}

int mcb_solver_support_vector_rows(mcb_solver *s, const int32_t *phases, int32_t count, uint64_t *out) {
    if (!s || count < 0 || (count > 0 && (!phases || !out))) return fail(MCB_E_INVALID, "invalid arguments");
    for (int32_t i = 0; i < count; ++i)
        if (phases[i] < 0 || phases[i] >= s->N) return fail(MCB_E_INVALID, "phase out of range");
    HIP_TRY(hipStreamSynchronize(s->stream));
    const size_t pitch = sizeof(uint64_t) * static_cast<size_t>(s->npad);
    for (int32_t i = 0; i < count; ++i) {
        const int32_t j = phases[i];
        uint64_t *row = out + static_cast<size_t>(i) * s->W;
        // column j of the word-major matrix: one word per row of W rows
        HIP_TRY(hipMemcpy2D(row, sizeof(uint64_t), s->Sw + j, pitch, sizeof(uint64_t), static_cast<size_t>(s->W),
                            hipMemcpyDeviceToHost));
        row[j >> 6] |= 1ull << (j & 63);
    }
    return MCB_OK;
}

int mcb_solver_cycle_vectors(mcb_solver *s, uint64_t *out, int32_t count) {
    if (!s || !out || count < 0 || count > s->N) return fail(MCB_E_INVALID, "invalid arguments");
    if (!s->C_store) return fail(MCB_E_STATE, "solver was created without store_cycles");
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipMemcpy(out, s->C_store, sizeof(uint64_t) * static_cast<size_t>(count) * s->W, hipMemcpyDeviceToHost));
    return MCB_OK;
}

int mcb_solver_parity(mcb_solver *s, const uint64_t *S, uint8_t *out) {
    if (!s || !S || !out) return fail(MCB_E_INVALID, "invalid arguments");
    if (s->lds_bytes == 0) return fail(MCB_E_STATE, "support vector too long for the per-phase parity kernel's LDS");
    uint64_t *dS = nullptr;
    uint8_t *dP = nullptr;
    int64_t dummy = 0;
    int rc = upload(&dS, S, s->W, dummy);
    if (rc) return rc;
    rc = dev_alloc(&dP, static_cast<size_t>(s->Ptot), dummy);
    if (rc) {
        (void)hipFree(dS);
        return rc;
    }
    PhaseArgs a = phase_args(s, 0);
    a.S = dS;
    a.parity_out = dP;
    a.wave_stats = nullptr;
    a.upd_tiles = 0;
    hipLaunchKernelGGL((k_phase<true, false>), dim3(s->blocks), dim3(s->wpb * kWave), s->lds_bytes, s->stream, a);
    hipError_t e1 = hipGetLastError();
    hipError_t e2 = hipStreamSynchronize(s->stream);
    hipError_t e3 = hipMemcpy(out, dP, static_cast<size_t>(s->Ptot), hipMemcpyDeviceToHost);
    (void)hipFree(dS);
    (void)hipFree(dP);
    if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) return fail(MCB_E_DEVICE, "parity kernel failed");
    return MCB_OK;
}

int mcb_solver_parity_noalloc(mcb_solver *s, const uint64_t *S, uint8_t *out) {
    if (!s || !S || !out) return fail(MCB_E_INVALID, "invalid arguments");
    if (s->lds_bytes == 0) return fail(MCB_E_STATE, "support vector too long for the per-phase parity kernel's LDS");
    const size_t s_bytes = (sizeof(uint64_t) * static_cast<size_t>(s->W) + 255) & ~static_cast<size_t>(255);
    if (s_bytes + static_cast<size_t>(s->Ptot) > kParityScratchBytes)
        return fail(MCB_E_INVALID, "support vector and tree positions exceed the " +
                                       std::to_string(kParityScratchBytes >> 20) +
                                       " MiB scratch of the no-allocation parity entry point");
    void *base = nullptr;
    int rc = scratch_address(&base);
    if (rc) return rc;
    uint64_t *dS = static_cast<uint64_t *>(base);
    uint8_t *dP = static_cast<uint8_t *>(base) + s_bytes;
    HIP_TRY(hipMemcpyAsync(dS, S, sizeof(uint64_t) * s->W, hipMemcpyHostToDevice, s->stream));
    PhaseArgs a = phase_args(s, 0);
    a.S = dS;
    a.parity_out = dP;  // the kernel writes every position of every root
    a.wave_stats = nullptr;
    a.upd_tiles = 0;
    hipLaunchKernelGGL((k_phase<true, false>), dim3(s->blocks), dim3(s->wpb * kWave), s->lds_bytes, s->stream, a);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, dP, static_cast<size_t>(s->Ptot), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return MCB_OK;
}

int mcb_solver_kernel_time(mcb_solver *s, int kernel, double *total_ms, int64_t *launches) {
    if (!s || kernel < 0 || kernel >= MCB_K_COUNT) return fail(MCB_E_INVALID, "invalid arguments");
    int rc = harvest_events(s);
    if (rc) return rc;
    if (total_ms) *total_ms = s->kt_ms[kernel];
    if (launches) *launches = s->kt_n[kernel];
    return MCB_OK;
}

int64_t mcb_solver_device_bytes(const mcb_solver *s) { return s ? s->bytes : -1; }

int mcb_solver_work_counters(mcb_solver *s, int64_t *positions, int64_t *candidates, int64_t *update_bytes) {
    if (!s) return fail(MCB_E_INVALID, "invalid arguments");
    HIP_TRY(hipStreamSynchronize(s->stream));
    std::vector<uint64_t> st(3 * static_cast<size_t>(s->blocks) * s->wpb);
    HIP_TRY(hipMemcpy(st.data(), s->wave_stats, sizeof(uint64_t) * st.size(), hipMemcpyDeviceToHost));
    uint64_t p = static_cast<uint64_t>(s->blk_positions), c = static_cast<uint64_t>(s->blk_candidates), u = 0;
    for (size_t i = 0; i < st.size(); i += 3) {
        p += st[i];
        c += st[i + 1];
        u += st[i + 2];
    }
    if (s->bk > 1) {
        unsigned long long ub = 0;
        unsigned long long slots[kByteSlots];
        HIP_TRY(hipMemcpy(slots, s->upd_bytes, sizeof(slots), hipMemcpyDeviceToHost));
        for (unsigned long long x : slots) ub += x;
        u += ub;
    }
    if (positions) *positions = static_cast<int64_t>(p);
    if (candidates) *candidates = static_cast<int64_t>(c);
    if (update_bytes) *update_bytes = static_cast<int64_t>(u);
    return MCB_OK;
}

int mcb_solver_launch_shape(const mcb_solver *s, int32_t *blocks, int32_t *waves_per_block, int32_t *lds_bytes,
                            int32_t *block_phases) {
    if (!s) return fail(MCB_E_INVALID, "invalid arguments");
    const bool blk = s->bk > 1;
    if (blocks) *blocks = blk ? s->q_blocks : s->blocks;
    if (waves_per_block) *waves_per_block = blk ? s->q_waves : s->wpb;
    if (lds_bytes) *lds_bytes = blk ? s->q_lds : s->lds_bytes;
    if (block_phases) *block_phases = s->bk;
    return MCB_OK;
}

int mcb_solver_update_grids(const mcb_solver *s, int32_t *dot_workgroups, int32_t *xor_workgroups) {
    if (!s) return fail(MCB_E_INVALID, "invalid arguments");
    if (dot_workgroups) *dot_workgroups = s->dot_wg;
    if (xor_workgroups) *xor_workgroups = s->xor_wg;
    return MCB_OK;
}

int mcb_hip_copy_bandwidth(int64_t bytes, int iters, double *gbps) {
    if (bytes <= 0 || iters <= 0 || !gbps) return fail(MCB_E_INVALID, "invalid arguments");
    int rc = ensure_device();
    if (rc) return rc;
    const int64_t n = bytes / 8;
    uint64_t *a = nullptr, *b = nullptr;
    HIP_TRY(hipMalloc(&a, n * 8));
    if (hipMalloc(&b, n * 8) != hipSuccess) {
        (void)hipFree(a);
        return fail(MCB_E_DEVICE, "hipMalloc");
    }
    (void)hipMemset(a, 1, n * 8);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    const unsigned grid = static_cast<unsigned>(g_dev.prop.multiProcessorCount) * 8;
    hipLaunchKernelGGL(k_copy_u64, dim3(grid), dim3(256), 0, nullptr, a, b, n);
    (void)hipEventRecord(e0, nullptr);
    for (int i = 0; i < iters; ++i) hipLaunchKernelGGL(k_copy_u64, dim3(grid), dim3(256), 0, nullptr, a, b, n);
    (void)hipEventRecord(e1, nullptr);
    hipError_t err = hipEventSynchronize(e1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    if (err != hipSuccess) return fail(MCB_E_DEVICE, "copy kernel failed");
    *gbps = 2.0 * static_cast<double>(n * 8) * iters / (ms * 1e-3) / 1e9;
    return MCB_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------------------------
// Batched form (host): the arena, its static launch tables, and one launch per kernel per step
// ---------------------------------------------------------------------------------------------

struct mcb_batch {
    mcb_solver_options opt{};
    int32_t G = 0, steps = 0, step = 0, merge_levels = 0;
    int64_t K = 0;                  // rank bound (ranks of all problems, each range padded to 64)
    std::vector<int32_t> order;     // problem at each arena position (phase count descending)
    std::vector<int32_t> slot_of;   // arena position of each problem
    std::vector<CompInfo> comps;    // by arena position
    std::vector<int64_t> rbase;     // rank base by arena position
    std::vector<int32_t> active;    // problems still running at each step
    std::vector<int32_t> chunk_cum, pairs_cum;                      // [G + 1] prefix counts
    std::vector<std::vector<int32_t>> grp_cum;                      // per merge level: [G + 1]
    std::vector<int64_t> grp_off;                                   // per merge level: first group
    std::vector<int64_t> step_off;                                  // per step: first entry of its tables
    std::vector<int32_t> dot_cum, xor_cum;
    // device arrays
    int64_t *tree_off = nullptr, *rc_off = nullptr;
    uint32_t *parent_pos = nullptr, *pos_pre = nullptr, *pos_end = nullptr, *rc_pu = nullptr, *rc_pv = nullptr,
             *rc_rank = nullptr, *cand_root = nullptr, *cand_pu = nullptr, *cand_pv = nullptr, *root_comp = nullptr;
    int32_t *pos_coord = nullptr, *rc_coord = nullptr, *cand_coord = nullptr;
    std::vector<TreePlan> plans;  // tree pass launch plans, ranges = problems
    int32_t plan = 0;
    uint64_t *q_scratch = nullptr;
    hipGraph_t graph = nullptr;   // every step, captured by mcb_batch_run_graph
    hipGraphExec_t graph_exec = nullptr;
    int64_t graph_positions = 0, graph_candidates = 0;
    CompInfo *d_comp = nullptr;
    ChunkInfo *d_chunks = nullptr;
    MergeGroup *d_groups = nullptr;
    int32_t *d_pairs_cum = nullptr, *d_dot_cum = nullptr, *d_xor_cum = nullptr;
    uint64_t *Sw = nullptr, *Tt = nullptr, *Qr = nullptr, *qbits = nullptr, *sum_q = nullptr, *blk_q = nullptr,
             *blk_m = nullptr, *blk_g = nullptr, *C_blk = nullptr, *Sf_blk = nullptr, *C_store = nullptr,
             *Ct = nullptr, *Bc = nullptr, *pair_c = nullptr;
    uint32_t *sum_cnt = nullptr, *sum_rank = nullptr, *sel = nullptr, *D_blk = nullptr, *Dcnt_blk = nullptr,
             *pw_list = nullptr, *pw_off = nullptr, *pw_cnt = nullptr, *pair_i = nullptr, *status = nullptr,
             *ulist = nullptr, *ucount = nullptr;
    unsigned long long *upd_bytes = nullptr;
    int64_t coords = 0, sw_words = 0, blk_words = 0, store_words = 0, slots = 0;
    std::vector<void *> allocs;
    hipStream_t stream = nullptr;
    int64_t bytes = 0;
    int64_t blk_positions = 0, blk_candidates = 0;
    // timing
    double kt_ms[MCB_K_COUNT] = {};
    int64_t kt_n[MCB_K_COUNT] = {};
    std::vector<hipEvent_t> ev_pool;
    std::vector<int> ev_kind;
    size_t ev_used = 0;
    bool timing_open = false;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;  // the open group's events, until launches take them
    bool timing_sample = true;  // this block (phase, step) records events
    int64_t timing_seq = 0;     // blocks (phases, steps) launched since reset
};

namespace {

template <typename T>
int bat_alloc(mcb_batch *b, T **p, size_t n) {
    int rc = dev_alloc(p, n, b->bytes);
    if (!rc) b->allocs.push_back(*p);
    return rc;
}

template <typename T, typename A>
int bat_upload(mcb_batch *b, T **p, const std::vector<T, A> &v) {
    int rc = upload(p, v.data(), v.size(), b->bytes);
    if (!rc) b->allocs.push_back(*p);
    return rc;
}

// Grid of k_batch_q: the plan's resident blocks rounded up to a multiple of kXcds, so that every
// XCD section has the same number of blocks and at least one.
unsigned batch_q_grid(const TreePlan &tp) {
    return static_cast<unsigned>((tp.blocks + kXcds - 1) / kXcds * kXcds);
}

BatchArgs batch_args(const mcb_batch *b, int32_t s) {
    BatchArgs a{};
    a.tree_off = b->tree_off;
    a.parent_pos = b->parent_pos;
    a.pos_coord = b->pos_coord;
    a.pos_pre = b->pos_pre;
    a.pos_end = b->pos_end;
    const TreePlan &tp = b->plans[b->plan];
    a.big_list = tp.big;
    a.small_list = tp.small;
    a.scratch = tp.scratch;
    a.huge_words = tp.huge_words;
    a.rc_off = b->rc_off;
    a.rc_pu = b->rc_pu;
    a.rc_pv = b->rc_pv;
    a.rc_rank = b->rc_rank;
    a.rc_coord = b->rc_coord;
    a.cand_root = b->cand_root;
    a.cand_pu = b->cand_pu;
    a.cand_pv = b->cand_pv;
    a.cand_coord = b->cand_coord;
    a.root_comp = b->root_comp;
    a.comp = b->d_comp;
    a.chunks = b->d_chunks;
    a.groups = b->d_groups;
    a.pairs_cum = b->d_pairs_cum;
    a.dot_cum = b->d_dot_cum + b->step_off[s];
    a.xor_cum = b->d_xor_cum + b->step_off[s];
    a.Sw = b->Sw;
    a.Tt = b->Tt;
    a.Qr = b->Qr;
    a.qbits = b->qbits;
    a.sum_q = b->sum_q;
    a.sum_cnt = b->sum_cnt;
    a.sum_rank = b->sum_rank;
    a.sel = b->sel;
    a.blk_q = b->blk_q;
    a.blk_m = b->blk_m;
    a.blk_g = b->blk_g;
    a.C_blk = b->C_blk;
    a.Sf_blk = b->Sf_blk;
    a.C_store = b->C_store;
    a.Ct = b->Ct;
    a.Bc = b->Bc;
    a.pair_c = b->pair_c;
    a.D_blk = b->D_blk;
    a.Dcnt_blk = b->Dcnt_blk;
    a.pw_list = b->pw_list;
    a.pw_off = b->pw_off;
    a.pw_cnt = b->pw_cnt;
    a.pair_i = b->pair_i;
    a.ulist = b->ulist;
    a.ucount = b->ucount;
    a.status = b->status;
    a.upd_bytes = b->upd_bytes;
    a.K = b->K;
    a.step = s;
    a.active = b->active[s];
    for (int x = 0; x < kXcds; ++x) {
        a.x_big_off[x] = tp.big_off[x];
        a.x_big_n[x] = tp.big_cum[x][a.active];
        a.x_small_off[x] = tp.small_off[x];
        a.x_small_n[x] = tp.small_cum[x][a.active];
    }
    a.big_words = tp.big_words;
    a.small_words = tp.small_words;
    a.xor_wt = (s + 1 + kTileDirect - 1) / kTileDirect;
    return a;
}

// Step s: block s of every problem with more than 64 s phases.
int launch_batch_step(mcb_batch *b, int32_t s) {
    BatchArgs a = batch_args(b, s);
    const int32_t A = a.active;
    const int64_t off = b->step_off[s];
    timing_sample(b);
    int rc = timed_begin(b, MCB_K_TRANSPOSE);
    if (rc) return rc;
    const int64_t tw = static_cast<int64_t>(A) * (s + 1);
    if ((rc = launch_on(b, true, k_batch_transpose, dim3(static_cast<unsigned>((tw + 3) / 4)), dim3(256), 0, a)))
        return rc;
    if ((rc = timed_end(b)) || (rc = timed_begin(b, MCB_K_PARITY_SEARCH))) return rc;
    const TreePlan &tp = b->plans[b->plan];
    if ((rc = launch_on(b, true, k_batch_q, dim3(batch_q_grid(tp)), dim3(tp.nw * kWave), tp.lds, a))) return rc;
    if ((rc = timed_end(b)) || (rc = timed_begin(b, MCB_K_SELECT))) return rc;
    // merge levels that some running problem has (no running problem has a level L+1 past the first empty one)
    int nm = 0;
    while (nm < b->merge_levels && b->grp_cum[nm][A] != 0) ++nm;
    a.n_items = b->chunk_cum[A];
    if ((rc = launch_on(b, nm == 0, k_batch_basis, dim3(static_cast<unsigned>((a.n_items + 3) / 4)), dim3(256), 0, a)))
        return rc;
    for (int L = 0; L < nm; ++L) {
        const int32_t n = b->grp_cum[L][A];
        a.groups = b->d_groups + b->grp_off[L];
        a.n_items = n;
        if ((rc = launch_on(b, L + 1 == nm, k_batch_merge, dim3(static_cast<unsigned>(n)), dim3(kMergeWaves * kWave),
                            0, a, L)))
            return rc;
    }
    if ((rc = timed_end(b)) || (rc = timed_begin(b, MCB_K_CYCLE))) return rc;
    // problems are in descending phase count, so problem 0 has the most words
    if ((rc = launch_on(b, false, k_batch_cycles, dim3(static_cast<unsigned>(A) * kBlockK), dim3(256),
                        sizeof(uint64_t) * b->comps[0].W, a)))
        return rc;
    if ((rc = launch_on(b, false, k_batch_final, dim3(static_cast<unsigned>((tw + 3) / 4)), dim3(256), 0, a)))
        return rc;
    if ((rc = launch_on(b, true, k_batch_pairs, dim3(static_cast<unsigned>(b->pairs_cum[A])), dim3(256),
                        sizeof(uint64_t) * s + 2 * 256 * sizeof(uint32_t), a)))
        return rc;
    if ((rc = timed_end(b)) || (rc = timed_begin(b, MCB_K_UPDATE))) return rc;
    const int32_t nd = b->dot_cum[off + A];
    if (nd > 0) {
        const size_t dlds = sizeof(uint64_t) * (kBlockK + kSegPairs) + sizeof(uint32_t) * (kSegPairs + 2 * kSegWords + 1);
        if ((rc = launch_on(b, false, k_batch_dot, dim3(static_cast<unsigned>(nd)), dim3(kDotWaves * kWave), dlds, a)))
            return rc;
    }
    if ((rc = launch_on(b, true, k_batch_xor, dim3(static_cast<unsigned>(b->xor_cum[off + A])), dim3(256),
                        sizeof(uint64_t) * kBlockK * kTileDirect, a)))
        return rc;
    b->blk_positions += tp.pos_cum[A];
    b->blk_candidates += tp.cand_cum[A];
    return timed_end(b);
}

}  // namespace

extern "C" {

void mcb_batch_destroy(mcb_batch *b) {
    if (!b) return;
    for (void *p : b->allocs)
        if (p) (void)hipFree(p);
    for (TreePlan &p : b->plans) free_plan(p);
    if (b->q_scratch) (void)hipFree(b->q_scratch);
    if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (hipEvent_t e : b->ev_pool) (void)hipEventDestroy(e);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

int mcb_batch_reset(mcb_batch *b) {
    if (!b) return fail(MCB_E_INVALID, "invalid arguments");
    HIP_TRY(hipStreamSynchronize(b->stream));
    HIP_TRY(hipMemsetAsync(b->Sw, 0, sizeof(uint64_t) * static_cast<size_t>(b->sw_words), b->stream));
    HIP_TRY(hipMemsetAsync(b->Tt, 0, sizeof(uint64_t) * static_cast<size_t>(b->coords), b->stream));
    HIP_TRY(hipMemsetAsync(b->qbits, 0, sizeof(uint64_t) * static_cast<size_t>(b->K / 64), b->stream));
    HIP_TRY(hipMemsetAsync(b->upd_bytes, 0, sizeof(unsigned long long) * kByteSlots, b->stream));
    HIP_TRY(hipMemsetAsync(b->status, 0, sizeof(uint32_t), b->stream));
    hipLaunchKernelGGL(k_fill_u32, dim3(static_cast<unsigned>((b->coords + 255) / 256)), dim3(256), 0, b->stream,
                       b->sel, kNoRank, b->coords);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(b->stream));
    b->step = 0;
    b->blk_positions = 0;
    b->blk_candidates = 0;
    b->ev_used = 0;
    b->timing_seq = 0;
    for (int i = 0; i < MCB_K_COUNT; ++i) {
        b->kt_ms[i] = 0;
        b->kt_n[i] = 0;
    }
    return MCB_OK;
}

int mcb_batch_create(const mcb_solver_desc *descs, int32_t count, const mcb_solver_options *opt, mcb_batch **out) {
    if (!descs || count <= 0 || !out) return fail(MCB_E_INVALID, "invalid arguments");
    SetupClock clock;
    *out = nullptr;
    int32_t largest = 0;
    {
        // every problem checked on the host threads; the first failure (lowest index) is reported
        std::vector<int32_t> mp(count, 0);
        std::vector<std::string> why(count);
        parallel_for(count, 1, [&](int64_t i) {
            if (validate_desc(&descs[i], &mp[i])) why[i] = g_err.empty() ? std::string("invalid") : g_err;
        });
        for (int32_t i = 0; i < count; ++i) {
            if (!why[i].empty()) return fail(MCB_E_INVALID, "problem " + std::to_string(i) + ": " + why[i]);
            largest = std::max(largest, mp[i]);
        }
    }
    if (opt && (opt->merge_group < 0 || opt->merge_group == 1 || opt->merge_group > 64))
        return fail(MCB_E_INVALID, "merge_group must be 0 or 2..64");
    if (opt && opt->timing_every < 0) return fail(MCB_E_INVALID, "timing_every must be 0 or more");
    if (opt && (opt->segment_resolve < 0 || opt->segment_resolve > 1))
        return fail(MCB_E_INVALID, "segment_resolve must be 0 or 1");
    int rc = ensure_device();
    if (rc) return rc;
    mcb_batch *b = new (std::nothrow) mcb_batch();
    if (!b) return fail(MCB_E_NOMEM, "out of memory");
    auto cleanup = [&](int code) {
        mcb_batch_destroy(b);
        return code;
    };
    if (opt) b->opt = *opt;
    else mcb_solver_options_init(&b->opt);
    const int32_t G = count;
    b->G = G;
    b->order.resize(G);
    std::iota(b->order.begin(), b->order.end(), 0);
    std::stable_sort(b->order.begin(), b->order.end(),
                     [&](int32_t x, int32_t y) { return descs[x].num_non_tree > descs[y].num_non_tree; });
    b->slot_of.resize(G);
    for (int32_t g = 0; g < G; ++g) b->slot_of[b->order[g]] = g;

    // Bases of every problem's ranges.
    int64_t F_tot = 0, P_tot = 0, K_tot = 0, rank_tot = 0, coords = 0, sw = 0, blk = 0, pw = 0, pwoff = 0, store = 0,
            slots = 0;
    b->comps.resize(G);
    b->rbase.resize(G);
    std::vector<int64_t> root_base(G), pos_base(G), cand_base(G), slot_base(G);
    std::vector<std::vector<int64_t>> lvl_n(G), lvl_slot(G);
    std::vector<int64_t> gsize(G, 16);
    int32_t max_levels = 0;
    for (int32_t g = 0; g < G; ++g) {
        const mcb_solver_desc &d = descs[b->order[g]];
        CompInfo &c = b->comps[g];
        c.N = d.num_non_tree;
        c.W = (c.N + 63) / 64;
        c.npad = static_cast<int64_t>(c.W) * 64;
        c.cb = coords;
        c.rb = rank_tot;
        c.sb = sw;
        c.cblk = blk;
        c.pw = pw;
        c.pwoff = pwoff;
        c.store = store;
        b->rbase[g] = rank_tot;
        root_base[g] = F_tot;
        pos_base[g] = P_tot;
        cand_base[g] = K_tot;
        slot_base[g] = slots;
        int64_t n = (d.num_candidates + kChunk - 1) / kChunk;
        gsize[g] = merge_group_size(n, b->opt.merge_group);
        for (;;) {
            lvl_slot[g].push_back(slots);
            lvl_n[g].push_back(n);
            slots += n;
            if (n == 1) break;
            n = (n + gsize[g] - 1) / gsize[g];
        }
        c.levels = static_cast<int32_t>(lvl_n[g].size());
        if (c.levels > kMaxLevels) return cleanup(fail(MCB_E_INVALID, "too many candidates for the basis levels"));
        max_levels = std::max(max_levels, c.levels);
        c.top = lvl_slot[g].back();
        F_tot += d.num_roots;
        P_tot += d.num_positions;
        K_tot += d.num_candidates;
        rank_tot += (d.num_candidates + 63) / 64 * 64;
        coords += c.npad;
        sw += static_cast<int64_t>(c.W) * c.npad;
        blk += static_cast<int64_t>(kBlockK) * c.W;
        pw += c.W;
        pwoff += c.W + 1;
        store += static_cast<int64_t>(c.N) * c.W;
    }
    if (rank_tot >= static_cast<int64_t>(kNoRank) || coords > 0x7FFFFFFF || F_tot > 0x7FFFFFFF)
        return cleanup(fail(MCB_E_INVALID, "the batch is too large for 32-bit ranks, coordinates or roots"));
    b->K = rank_tot;
    b->coords = coords;
    b->sw_words = sw;
    b->blk_words = blk;
    b->store_words = store;
    b->slots = slots;
    b->merge_levels = max_levels - 1;
    b->steps = (b->comps[0].N + kBlockK - 1) / kBlockK;
    // the largest problem's cycle vectors and pair list stage in LDS (as in the single-problem form)
    if ((rc = reserve_dynamic_lds(reinterpret_cast<const void *>(k_batch_cycles), 8ll * b->comps[0].W,
                                  "support vector too long for the cycle kernel's LDS")) ||
        (rc = reserve_dynamic_lds(reinterpret_cast<const void *>(k_batch_pairs), 8ll * b->steps + 2048,
                                  "support vector too long for the pair kernel's LDS")))
        return cleanup(rc);

    // Arena arrays: trees and candidates, problem after problem in arena order, coordinates and ranks
    // shifted into each problem's range.
    RawVec<int64_t> tree_off(F_tot + 1), rc_off(F_tot + 1);
    RawVec<uint32_t> parent_pos(P_tot), rc_pu(K_tot), rc_pv(K_tot), rc_rank(K_tot), cand_root(rank_tot),
        cand_pu(rank_tot), cand_pv(rank_tot), root_comp(F_tot);
    RawVec<int32_t> pos_coord(P_tot), rc_coord(K_tot), cand_coord(rank_tot);
    std::vector<int64_t> ranges(G + 1, F_tot);
    clock.mark("bases and tables");
    parallel_for(G, 1, [&](int64_t gi) {
        const int32_t g = static_cast<int32_t>(gi);
        const mcb_solver_desc &d = descs[b->order[g]];
        const CompInfo &c = b->comps[g];
        const int64_t R = root_base[g], PB = pos_base[g], KB = cand_base[g], rb = c.rb, cb = c.cb;
        for (int32_t r = 0; r < d.num_roots; ++r) {
            tree_off[R + r] = PB + d.tree_offset[r];
            rc_off[R + r] = KB + d.root_cand_offset[r];
            root_comp[R + r] = static_cast<uint32_t>(g);
        }
        for (int64_t i = 0; i < d.num_positions; ++i) {
            parent_pos[PB + i] = d.parent_pos[i];
            pos_coord[PB + i] = d.pos_coord[i] >= 0 ? static_cast<int32_t>(cb + d.pos_coord[i]) : -1;
        }
        for (int64_t k = 0; k < d.num_candidates; ++k) {
            cand_root[rb + k] = static_cast<uint32_t>(R + d.cand_root[k]);
            cand_pu[rb + k] = d.cand_pos_u[k];
            cand_pv[rb + k] = d.cand_pos_v[k];
            cand_coord[rb + k] = d.cand_coord[k] >= 0 ? static_cast<int32_t>(cb + d.cand_coord[k]) : -1;
        }
        // the padding of the problem's rank range (up to the next range, a multiple of 64) holds no candidate
        const int64_t rank_end = g + 1 < G ? b->comps[g + 1].rb : rank_tot;
        for (int64_t k = rb + d.num_candidates; k < rank_end; ++k) {
            cand_root[k] = cand_pu[k] = cand_pv[k] = 0;
            cand_coord[k] = -1;
        }
        for (int64_t i = 0; i < d.num_candidates; ++i) {
            const uint32_t k = d.root_cand_rank[i];
            rc_pu[KB + i] = d.cand_pos_u[k];
            rc_pv[KB + i] = d.cand_pos_v[k];
            rc_rank[KB + i] = static_cast<uint32_t>(rb + k);
            rc_coord[KB + i] = d.cand_coord[k] >= 0 ? static_cast<int32_t>(cb + d.cand_coord[k]) : -1;
        }
        ranges[g] = R;
    });
    tree_off[F_tot] = P_tot;
    rc_off[F_tot] = K_tot;
    clock.mark("arena assembly");
    RawVec<uint32_t> pre, end;
    preorder_intervals(tree_off.data(), parent_pos.data(), static_cast<int32_t>(F_tot), pre, end);
    clock.mark("preorder intervals");

    if ((rc = make_plans(b->opt.waves_per_block, b->opt.tree_cap, kXcds, tree_off.data(), parent_pos.data(),
                         rc_off.data(), ranges,
                         reinterpret_cast<const void *>(k_batch_q), b->plans, b->bytes, &b->q_scratch)))
        return cleanup(rc);
    if (b->plans.empty()) return cleanup(fail(MCB_E_INVALID, "a tree is too large for the blocked form's LDS budget"));
    clock.mark("tree pass plans");

    // Static launch tables: first-level chunks, merge groups per level, pairs workgroups.
    std::vector<ChunkInfo> chunks;
    b->chunk_cum.assign(G + 1, 0);
    b->pairs_cum.assign(G + 1, 0);
    for (int32_t g = 0; g < G; ++g) {
        const CompInfo &c = b->comps[g];
        const int64_t words = (descs[b->order[g]].num_candidates + 63) / 64;
        for (int64_t i = 0; i < lvl_n[g][0]; ++i) {
            ChunkInfo ch{};
            ch.r0 = c.rb + i * kChunk;
            ch.slot = lvl_slot[g][0] + i;
            ch.comp = g;
            ch.words = static_cast<int32_t>(std::min<int64_t>(kWave, words - i * kWave));
            chunks.push_back(ch);
        }
        b->chunk_cum[g + 1] = b->chunk_cum[g] + static_cast<int32_t>(lvl_n[g][0]);
        b->pairs_cum[g + 1] = b->pairs_cum[g] + 1 + (c.W + 3) / 4;
    }
    std::vector<MergeGroup> groups;
    b->grp_cum.assign(std::max(0, b->merge_levels), std::vector<int32_t>(G + 1, 0));
    b->grp_off.assign(std::max(0, b->merge_levels), 0);
    for (int L = 0; L < b->merge_levels; ++L) {
        b->grp_off[L] = static_cast<int64_t>(groups.size());
        for (int32_t g = 0; g < G; ++g) {
            int32_t n = 0;
            if (b->comps[g].levels >= L + 2) {
                n = static_cast<int32_t>(lvl_n[g][L + 1]);
                for (int32_t gi = 0; gi < n; ++gi) {
                    MergeGroup mg{};
                    mg.src = lvl_slot[g][L] + static_cast<int64_t>(gi) * gsize[g];
                    mg.dst = lvl_slot[g][L + 1] + gi;
                    mg.comp = g;
                    mg.members =
                        static_cast<int32_t>(std::min<int64_t>(gsize[g], lvl_n[g][L] - static_cast<int64_t>(gi) * gsize[g]));
                    groups.push_back(mg);
                }
            }
            b->grp_cum[L][g + 1] = b->grp_cum[L][g] + n;
        }
    }
    // Per step: running problems and the prefix counts of their dot and xor workgroups.
    b->active.resize(b->steps);
    b->step_off.resize(b->steps);
    for (int32_t s = 0; s < b->steps; ++s) {
        int32_t A = 0;
        while (A < G && b->comps[A].N > s * kBlockK) ++A;
        b->active[s] = A;
        b->step_off[s] = static_cast<int64_t>(b->dot_cum.size());
        const int64_t wt = (s + 1 + kTileDirect - 1) / kTileDirect;
        int64_t dsum = 0, xsum = 0;
        b->dot_cum.push_back(0);
        b->xor_cum.push_back(0);
        for (int32_t g = 0; g < A; ++g) {
            const int64_t left = b->comps[g].N - static_cast<int64_t>(s) * kBlockK;
            const int64_t later = left - std::min<int64_t>(left, kBlockK);
            dsum += (later + kDotCols - 1) / kDotCols;
            xsum += (left + 255) / 256 * wt;
            if (dsum > 0x7FFFFFFF || xsum > 0x7FFFFFFF) return cleanup(fail(MCB_E_INVALID, "batch too large"));
            b->dot_cum.push_back(static_cast<int32_t>(dsum));
            b->xor_cum.push_back(static_cast<int32_t>(xsum));
        }
    }

    std::vector<CompInfo> comps = b->comps;
    clock.mark("launch tables");
    if ((rc = bat_upload(b, &b->tree_off, tree_off)) || (rc = bat_upload(b, &b->parent_pos, parent_pos)) ||
        (rc = bat_upload(b, &b->pos_coord, pos_coord)) || (rc = bat_upload(b, &b->pos_pre, pre)) ||
        (rc = bat_upload(b, &b->pos_end, end)) || (rc = bat_upload(b, &b->rc_off, rc_off)) || (rc = bat_upload(b, &b->rc_pu, rc_pu)) ||
        (rc = bat_upload(b, &b->rc_pv, rc_pv)) || (rc = bat_upload(b, &b->rc_rank, rc_rank)) ||
        (rc = bat_upload(b, &b->rc_coord, rc_coord)) || (rc = bat_upload(b, &b->cand_root, cand_root)) ||
        (rc = bat_upload(b, &b->cand_pu, cand_pu)) || (rc = bat_upload(b, &b->cand_pv, cand_pv)) ||
        (rc = bat_upload(b, &b->cand_coord, cand_coord)) || (rc = bat_upload(b, &b->root_comp, root_comp)) ||
        (rc = bat_upload(b, &b->d_comp, comps)) || (rc = bat_upload(b, &b->d_chunks, chunks)) ||
        (rc = bat_upload(b, &b->d_groups, groups)) || (rc = bat_upload(b, &b->d_pairs_cum, b->pairs_cum)) ||
        (rc = bat_upload(b, &b->d_dot_cum, b->dot_cum)) || (rc = bat_upload(b, &b->d_xor_cum, b->xor_cum)))
        return cleanup(rc);
    if ((rc = bat_alloc(b, &b->Sw, sw)) || (rc = bat_alloc(b, &b->Tt, coords)) || (rc = bat_alloc(b, &b->Ct, coords)) ||
        (rc = bat_alloc(b, &b->Bc, coords)) || (rc = bat_alloc(b, &b->sel, coords)) ||
        (rc = bat_alloc(b, &b->Qr, rank_tot)) || (rc = bat_alloc(b, &b->qbits, rank_tot / 64)) ||
        (rc = bat_alloc(b, &b->sum_cnt, slots)) || (rc = bat_alloc(b, &b->sum_rank, slots * kWave)) ||
        (rc = bat_alloc(b, &b->sum_q, slots * kWave)) || (rc = bat_alloc(b, &b->blk_q, G * kBlockK)) ||
        (rc = bat_alloc(b, &b->blk_m, G * kBlockK)) || (rc = bat_alloc(b, &b->blk_g, G * kBlockK)) ||
        (rc = bat_alloc(b, &b->Dcnt_blk, G * kBlockK)) || (rc = bat_alloc(b, &b->C_blk, blk)) ||
        (rc = bat_alloc(b, &b->Sf_blk, blk)) || (rc = bat_alloc(b, &b->D_blk, blk)) ||
        (rc = bat_alloc(b, &b->pair_i, blk)) || (rc = bat_alloc(b, &b->pair_c, blk)) ||
        (rc = bat_alloc(b, &b->pw_list, pw)) || (rc = bat_alloc(b, &b->pw_off, pwoff)) ||
        (rc = bat_alloc(b, &b->pw_cnt, G)) || (rc = bat_alloc(b, &b->status, 1)) ||
        (rc = bat_alloc(b, &b->ulist, coords)) || (rc = bat_alloc(b, &b->ucount, G)) ||
        (rc = bat_alloc(b, &b->upd_bytes, kByteSlots)))
        return cleanup(rc);
    if (b->opt.store_cycles && (rc = bat_alloc(b, &b->C_store, store))) return cleanup(rc);
    clock.mark("uploads and allocations");
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
        return cleanup(fail(MCB_E_DEVICE, "hipStreamCreate"));
    if (b->opt.timing) {
        b->ev_pool.resize(4096);
        b->ev_kind.resize(2048);
        for (auto &e : b->ev_pool)
            if (hipEventCreate(&e) != hipSuccess) return cleanup(fail(MCB_E_DEVICE, "hipEventCreate"));
    }
    if ((rc = mcb_batch_reset(b))) return cleanup(rc);
    if (b->plans.size() > 1) {
        // tree pass width: time every plan on step 0 (every problem, from the unit support vectors)
        const BatchArgs a = batch_args(b, 0);
        hipLaunchKernelGGL(k_batch_transpose, dim3(static_cast<unsigned>((a.active + 3) / 4)), dim3(256), 0, b->stream,
                           a);
        if (hipGetLastError() != hipSuccess) return cleanup(fail(MCB_E_DEVICE, "k_batch_transpose"));
        rc = pick_fastest(
            b->plans.size(), b->stream, [&](size_t i) { b->plan = static_cast<int32_t>(i); },
            [&]() -> int {
                const BatchArgs q = batch_args(b, 0);
                const TreePlan &tp = b->plans[b->plan];
                hipLaunchKernelGGL(k_batch_q, dim3(batch_q_grid(tp)), dim3(tp.nw * kWave), tp.lds, b->stream, q);
                HIP_TRY(hipGetLastError());
                return MCB_OK;
            });
        if (rc || (rc = mcb_batch_reset(b))) return cleanup(rc);
    }
    clock.mark("reset and plan timing");
    *out = b;
    return MCB_OK;
}

int32_t mcb_batch_steps(const mcb_batch *b) { return b ? b->steps : -1; }

int32_t mcb_batch_step(const mcb_batch *b) { return b ? b->step : -1; }

int mcb_batch_run(mcb_batch *b, int32_t steps) {
    if (!b || steps < 0) return fail(MCB_E_INVALID, "invalid arguments");
    const int32_t end = std::min(b->steps, b->step + steps);
    while (b->step < end) {
        int rc = launch_batch_step(b, b->step);
        if (rc) return rc;
        ++b->step;
    }
    return MCB_OK;
}

int mcb_batch_run_graph(mcb_batch *b) {
    if (!b) return fail(MCB_E_INVALID, "invalid arguments");
    if (b->step != 0) return fail(MCB_E_STATE, "the graph runs every step from step 0 (reset first)");
    if (b->opt.timing) return fail(MCB_E_STATE, "graph replay needs timing off");
    if (!b->graph_exec) {
        HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        int rc = MCB_OK;
        for (int32_t st = 0; st < b->steps && rc == MCB_OK; ++st) rc = launch_batch_step(b, st);
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(b->stream, &g);
        if (rc || ec != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return rc ? rc : fail(MCB_E_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
        }
        b->graph = g;
        HIP_TRY(hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0));
        // the counters were zero at phase 0; capture ran the host side of every launch once
        b->graph_positions = b->blk_positions;
        b->graph_candidates = b->blk_candidates;
    }
    HIP_TRY(hipGraphLaunch(b->graph_exec, b->stream));
    b->blk_positions = b->graph_positions;
    b->blk_candidates = b->graph_candidates;
    b->step = b->steps;
    return MCB_OK;
}

int mcb_batch_sync(mcb_batch *b) {
    if (!b) return fail(MCB_E_INVALID, "invalid arguments");
    HIP_TRY(hipStreamSynchronize(b->stream));
    int rc = harvest_events(b);
    if (rc) return rc;
    uint32_t st = 0;
    HIP_TRY(hipMemcpy(&st, b->status, sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (st) return fail(MCB_E_NOT_FOUND, "a phase found no odd candidate");
    return MCB_OK;
}

int mcb_batch_selected(mcb_batch *b, int32_t problem, int32_t *out, int32_t count) {
    if (!b || !out || problem < 0 || problem >= b->G) return fail(MCB_E_INVALID, "invalid arguments");
    const int32_t g = b->slot_of[problem];
    const CompInfo &c = b->comps[g];
    if (count < 0 || count > c.N) return fail(MCB_E_INVALID, "invalid arguments");
    HIP_TRY(hipStreamSynchronize(b->stream));
    std::vector<uint32_t> v(static_cast<size_t>(count));
    if (count) HIP_TRY(hipMemcpy(v.data(), b->sel + c.cb, sizeof(uint32_t) * count, hipMemcpyDeviceToHost));
    for (int32_t i = 0; i < count; ++i)
        out[i] = v[i] == kNoRank ? -1 : static_cast<int32_t>(static_cast<int64_t>(v[i]) - b->rbase[g]);
    return MCB_OK;
}

int mcb_batch_cycle_vectors(mcb_batch *b, int32_t problem, uint64_t *out, int32_t count) {
    if (!b || !out || problem < 0 || problem >= b->G) return fail(MCB_E_INVALID, "invalid arguments");
    if (!b->C_store) return fail(MCB_E_STATE, "batch was created without store_cycles");
    const CompInfo &c = b->comps[b->slot_of[problem]];
    if (count < 0 || count > c.N) return fail(MCB_E_INVALID, "invalid arguments");
    HIP_TRY(hipStreamSynchronize(b->stream));
    if (count)
        HIP_TRY(hipMemcpy(out, b->C_store + c.store, sizeof(uint64_t) * static_cast<size_t>(count) * c.W,
                          hipMemcpyDeviceToHost));
    return MCB_OK;
}

int mcb_batch_kernel_time(mcb_batch *b, int kernel, double *total_ms, int64_t *launches) {
    if (!b || kernel < 0 || kernel >= MCB_K_COUNT) return fail(MCB_E_INVALID, "invalid arguments");
    int rc = harvest_events(b);
    if (rc) return rc;
    if (total_ms) *total_ms = b->kt_ms[kernel];
    if (launches) *launches = b->kt_n[kernel];
    return MCB_OK;
}

int64_t mcb_batch_device_bytes(const mcb_batch *b) { return b ? b->bytes : -1; }

int mcb_batch_work_counters(mcb_batch *b, int64_t *positions, int64_t *candidates, int64_t *update_bytes) {
    if (!b) return fail(MCB_E_INVALID, "invalid arguments");
    HIP_TRY(hipStreamSynchronize(b->stream));
    unsigned long long ub = 0;
    unsigned long long slots[kByteSlots];
    HIP_TRY(hipMemcpy(slots, b->upd_bytes, sizeof(slots), hipMemcpyDeviceToHost));
    for (unsigned long long x : slots) ub += x;
    if (positions) *positions = b->blk_positions;
    if (candidates) *candidates = b->blk_candidates;
    if (update_bytes) *update_bytes = static_cast<int64_t>(ub);
    return MCB_OK;
}

int mcb_batch_launch_shape(const mcb_batch *b, int32_t *blocks, int32_t *waves_per_block, int32_t *lds_bytes,
                           int32_t *steps) {
    if (!b) return fail(MCB_E_INVALID, "invalid arguments");
    const TreePlan &tp = b->plans[b->plan];
    if (blocks) *blocks = static_cast<int32_t>(batch_q_grid(tp));
    if (waves_per_block) *waves_per_block = tp.nw;
    if (lds_bytes) *lds_bytes = tp.lds;
    if (steps) *steps = b->steps;
    return MCB_OK;
}

}  // extern "C"
