// gfx950 shortest-path trees of the FVS roots: the stopped searches of the host preprocessing
// (include/Dijkstra.h:78-127 with the candidate filter of include/Dijkstra.h:205-254), one wave per
// root, every root of a problem in one persistent launch.
//
// Exactness.  Each root's tree must be the one the reference's std::priority_queue search builds:
// among equal distances the heap's own pop order decides parents, and so the candidate set.  The
// search is therefore the sequential one, heap moves included.  What the wave adds is width inside
// each step:
//   * the heap (vertex, distance, row start) lives in LDS; a pop finds the hole's path six levels per
//     round from one load of every node's two children (a ballot of "left child strictly smaller",
//     then a walk over the ballot bits), and a push finds its sift-up depth from one load of all
//     ancestors (their distances fall monotonically towards the top, so "parent larger than the new
//     key" is a prefix, counted by a ballot).  Every move is then one LDS store per lane, and the
//     result is the element order of libstdc++'s std::push_heap / std::pop_heap;
//   * the popped vertex's row is relaxed with one lane per edge, in row order: parallel edges to one
//     vertex sit next to each other (rows are sorted by column), and a row holding such a run is
//     relaxed lane by lane so that every strict improvement pushes, as the sequential loop does.
// Per-vertex state is global memory in a per-wave slab, stamped with the root's epoch so that a new
// root starts without clearing.  The search stops once every reached, unfinalised vertex is blocked
// (the pops up to then are the full search's; mcb_host.cpp's dijkstra does the same).  The wave then
// lists the accepted candidate edges (first direction, both ends finalised and unblocked, not a tree
// edge, meeting first at the root) and the vertices of the pruned layout (candidate endpoints and their
// ancestors, marked by atomic exchange), and copies both into pools reserved by one atomic per root.
// The host orders them (mcb_problem_finish_trees): the reference's candidate grouping and rank, and
// the BFS-level layout.
//
// A root whose heap outgrows the LDS capacity of a stage is searched again in the next stage with a
// larger heap and fewer waves per CU; one that outgrows the last stage is left to the host library.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "mcb_hip.h"

namespace {

thread_local std::string g_sp_err;

int sfail(int code, const std::string &msg) {
    g_sp_err = msg;
    return code;
}

#define SP_TRY(call)                                                                             \
    do {                                                                                         \
        hipError_t _e = (call);                                                                  \
        if (_e != hipSuccess)                                                                    \
            return sfail(MCB_E_DEVICE, std::string(#call) + ": " + hipGetErrorString(_e));      \
    } while (0)

constexpr uint32_t kIn = 1u;          // state.w bit 0: finalised
constexpr uint32_t kBlk = 2u;         // state.w bit 1: blocked (tentative until finalised); level in bits 2..31
constexpr uint32_t kRootBit = 0x80000000u;  // colr: the far end is an FVS root
constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr int kStages = 3;
constexpr int kMaxWavesPerCu = 16;  // workspaces are made for this many waves per CU

struct SpArgs {
    const uint32_t *ro;     // [n+1] row offsets
    const uint32_t *colr;   // [2m] far end | kRootBit when it is a root
    const int32_t *w;       // [2m]
    const uint32_t *rs;     // [2m] row start of the far end
    const uint32_t *rows;   // [2m] near end
    const uint32_t *rev;    // [2m] opposite direction
    const uint32_t *roots;  // [F] vertex of each root index
    const uint32_t *work;   // root indices of this launch
    uint32_t n, m2, num_work, heap_cap, epoch_base;
    uint32_t *next;         // work counter
    // per-wave slabs (slab = workgroup index)
    uint4 *state;           // [slabs][n] (epoch, distance, parent edge, bits)
    uint2 *aux;             // [slabs][n] (branch: the root's child above the vertex, finalisation index)
    uint32_t *keep;         // [slabs][n] epoch of the layout mark
    uint32_t *open_list;    // [slabs][n] finalised unblocked vertices
    uint2 *cand_buf;        // [slabs][cand_stride] (edge, weight)
    uint4 *kept_buf;        // [slabs][n] (vertex, parent edge, level, finalisation index)
    uint64_t cand_stride;
    // per root
    int32_t *status;        // 0 done, 1 heap outgrew the stage, 2 pools full
    int64_t *cand_start;
    int32_t *cand_count;
    int64_t *kept_start;
    int32_t *kept_count;
    unsigned long long *pool_top;  // [2] reserved candidates, kept vertices (counts past the capacity)
    uint32_t *cand_edge_pool;
    int32_t *cand_weight_pool;
    uint4 *kept_pool;
    uint64_t cand_cap, kept_cap;
};

__device__ __forceinline__ uint32_t uni(uint32_t x) {
    return static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(x)));
}
__device__ __forceinline__ uint64_t uni64(uint64_t x) {
    return (static_cast<uint64_t>(uni(static_cast<uint32_t>(x >> 32))) << 32) | uni(static_cast<uint32_t>(x));
}
__device__ __forceinline__ uint32_t lane_u32(uint32_t x, int l) {
    return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(x), l));
}
// LDS written by one lane and read by another: the wave's LDS operations complete in order, so only
// the compiler must not move them across this point.
__device__ __forceinline__ void lds_order() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}
// Global memory written by one lane and read by another lane of the same wave.
__device__ __forceinline__ void mem_order() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

// Heap entries: (distance, vertex) in hp, the vertex's row start in hr; comparisons use the distance
// only.  The caller keeps the top entry in scalar registers (td, tv, tr): every step reads it, and both
// operations know what they leave at index 0.

// std::push_heap of (vd, vv, vr): the new element at index len sifts up while its parent's distance
// is strictly larger.  Returns false (heap unchanged) when the heap is full.
__device__ __forceinline__ bool heap_push(uint2 *hp, uint32_t *hr, uint32_t &len, uint32_t cap, int32_t vd,
                                          uint32_t vv, uint32_t vr, int32_t &td, uint32_t &tv, uint32_t &tr,
                                          int lane) {
    if (len >= cap) return false;
    const uint32_t hole = len;
    len = hole + 1;
    const uint32_t depth = 31u - static_cast<uint32_t>(__clz(hole + 1));  // ancestors of the hole
    const uint32_t k = static_cast<uint32_t>(lane) + 1;                 // lane k-1 looks at ancestor k
    const bool act = k <= depth;
    const uint32_t ak = act ? ((hole + 1) >> k) - 1 : 0u;
    const uint2 ek = hp[ak];
    const uint32_t rk = hr[ak];
    const uint32_t K = static_cast<uint32_t>(__popcll(__ballot(act && static_cast<int32_t>(ek.x) > vd)));
    lds_order();
    if (k <= K) {
        const uint32_t dst = ((hole + 1) >> (k - 1)) - 1;
        hp[dst] = ek;
        hr[dst] = rk;
    }
    const uint32_t fin = ((hole + 1) >> K) - 1;
    if (lane == 0) {
        hp[fin] = make_uint2(static_cast<uint32_t>(vd), vv);
        hr[fin] = vr;
    }
    lds_order();
    if (fin == 0) {
        td = vd;
        tv = vv;
        tr = vr;
    }
    return true;
}

// std::pop_heap of the top (td, tv, tr): with lp = len - 1 elements left, the hole walks from the top
// while its index is below (lp - 1) / 2, each step into the right child unless that one's distance is
// strictly larger than the left one's, then into a lone last child when lp is even; the old last element
// then sifts up from the hole.  The walk takes six levels per round: lane j loads both children of node j
// of the six-level subtree under the round's first hole (the top subtree's addresses are fixed, so the
// first round's loads go out with the last element's) and keeps the child it would step into; a ballot
// of those choices gives the path, and one lane permute per round hands lane i the entry the hole takes
// from below at path node p_i.  Leaves (td, tv, tr) the new top when elements remain.
__device__ __forceinline__ void heap_pop(uint2 *hp, uint32_t *hr, uint32_t &len, int32_t &td, uint32_t &tv,
                                         uint32_t &tr, int lane) {
    const uint32_t lp = len - 1;
    len = lp;
    if (lp == 0) return;
    const uint32_t j = static_cast<uint32_t>(lane);               // node j of a six-level subtree
    const uint32_t t = 31u - static_cast<uint32_t>(__clz(j + 1));  // its depth in the subtree
    const uint32_t lim = (lp - 1) >> 1;
    const uint2 last = hp[lp];
    const uint32_t last_r = hr[lp];
    uint32_t hole = 0, L = 0, my_p = 0;
    uint2 nxt = make_uint2(0u, 0u);  // lane i < L: the entry below path node p_i that moves up into it
    uint32_t nxt_r = 0;
    bool walking = lim > 0;
    while (walking) {
        const uint32_t g = ((hole + 1) << t) - 1 + (j + 1 - (1u << t));
        const bool valid = j < 63 && g < lim;
        const uint32_t c2 = valid ? 2 * g + 2 : 0u;
        const uint2 er = hp[c2];
        const uint2 el = hp[valid ? c2 - 1 : 0u];
        const uint32_t rr = hr[c2];
        const uint32_t rl = hr[valid ? c2 - 1 : 0u];
        const bool left = valid && static_cast<int32_t>(er.x) > static_cast<int32_t>(el.x);
        const uint64_t goes_left = __ballot(left);
        const uint64_t vmask = __ballot(valid);
        const uint32_t L0 = L;
        uint32_t loc = 0, src = 0;
        for (int s = 0; s < 6; ++s) {
            if (!((vmask >> loc) & 1ull)) {
                walking = false;
                break;
            }
            const uint32_t lb = static_cast<uint32_t>((goes_left >> loc) & 1ull);
            src = (j == L) ? loc : src;
            hole = 2 * hole + 2 - lb;
            ++L;
            my_p = (j == L) ? hole : my_p;
            loc = 2 * loc + 2 - lb;
        }
        const uint32_t cx = left ? el.x : er.x, cy = left ? el.y : er.y, cr = left ? rl : rr;
        const uint32_t fx = __shfl(cx, static_cast<int>(src)), fy = __shfl(cy, static_cast<int>(src));
        const uint32_t fr = __shfl(cr, static_cast<int>(src));
        const bool mine = j >= L0 && j < L;
        nxt = mine ? make_uint2(fx, fy) : nxt;
        nxt_r = mine ? fr : nxt_r;
    }
    if ((lp & 1u) == 0 && hole == ((lp - 2) >> 1)) {
        const uint2 lc = hp[2 * hole + 1];
        const uint32_t lcr = hr[2 * hole + 1];
        nxt = (j == L) ? lc : nxt;
        nxt_r = (j == L) ? lcr : nxt_r;
        hole = 2 * hole + 1;
        ++L;
        my_p = (j == L) ? hole : my_p;
    }
    const int32_t vd = static_cast<int32_t>(uni(last.x));
    const uint32_t K = static_cast<uint32_t>(__popcll(__ballot(j < L && static_cast<int32_t>(nxt.x) > vd)));
    const uint32_t cut = L - K;
    lds_order();
    if (j < cut) {
        hp[my_p] = nxt;
        hr[my_p] = nxt_r;
    } else if (j == cut) {
        hp[my_p] = last;
        hr[my_p] = last_r;
    }
    lds_order();
    if (cut > 0) {
        td = static_cast<int32_t>(lane_u32(nxt.x, 0));
        tv = lane_u32(nxt.y, 0);
        tr = lane_u32(nxt_r, 0);
    } else {
        td = vd;
        tv = uni(last.y);
        tr = uni(last_r);
    }
}


// One root: the stopped search, then its candidates and layout vertices.  Returns the status.
__device__ int search_root(const SpArgs &a, uint32_t r, uint2 *hp, uint32_t *hr, uint32_t *cnt,
                           uint4 *st, uint2 *ax, uint32_t *kp, uint32_t *ol, uint2 *cb, uint4 *kb, int lane) {
    const uint32_t src = uni(a.roots[r]);
    const uint32_t epoch = a.epoch_base + r + 1;
    const uint32_t H = a.heap_cap;
    if (lane == 0) {
        st[src] = make_uint4(epoch, 0u, kNone, 0u);
        ax[src] = make_uint2(kNone, 0u);
        hp[0] = make_uint2(0u, src);
        hr[0] = a.ro[src];
    }
    int32_t td = 0;
    uint32_t tv = src, tr = uni(a.ro[src]);
    mem_order();
    lds_order();
    uint32_t len = 1, nfin = 0, nopen = 0;
    int32_t open = 1;
    bool full = false;
    // The state and row of the next heap top are loaded with each step's neighbour loads (pf_*).  They
    // stand for that vertex when it is popped next, with the writes of the step in between forwarded
    // from the registers of the lanes that made them (fw_*): the prefetch was issued before them.
    bool pf_ok = false;
    uint32_t pf_v = kNone, pf_bits = 0, pf_br = 0, pf_rend = 0, pf_cr = 0, pf_rs = 0;
    int32_t pf_w = 0;
    uint64_t fw_mask = 0;
    uint32_t fw_c = kNone, fw_bits = 0, fw_br = 0, fw_fin = kNone;
    while (len > 0) {
        const uint32_t u = tv;
        const int32_t du = td;
        const uint32_t urow = tr;
        heap_pop(hp, hr, len, td, tv, tr, lane);
        const uint32_t nu = len > 0 ? tv : kNone;  // the next top unless this step pushes above it
        const uint32_t nrow = len > 0 ? tr : 0u;
        uint32_t sbits, bru, rend, cr0, rs0;
        int32_t w0;
        if (pf_ok && pf_v == u) {
            const uint64_t m = __ballot(((fw_mask >> lane) & 1ull) && fw_c == u);
            if (u == fw_fin) {
                sbits = kIn;
                bru = 0;
            } else if (m) {
                const int l = __ffsll(static_cast<unsigned long long>(m)) - 1;
                sbits = lane_u32(fw_bits, l);
                bru = lane_u32(fw_br, l);
            } else {
                sbits = pf_bits;
                bru = pf_br;
            }
            rend = pf_rend;
            cr0 = pf_cr;
            w0 = pf_w;
            rs0 = pf_rs;
        } else {
            mem_order();
            // the vertex's state and its row in one round trip (the row loads are speculative)
            const uint4 su4 = st[u];
            bru = uni(ax[u].x);
            rend = uni(a.ro[u + 1]);
            const uint32_t off0 = min(urow + static_cast<uint32_t>(lane), a.m2 - 1);
            cr0 = a.colr[off0];
            w0 = a.w[off0];
            rs0 = a.rs[off0];
            sbits = uni(su4.w);
        }
        pf_ok = false;
        if (sbits & kIn) continue;  // a stale entry (lazy deletion)
        const bool bu = (sbits & kBlk) != 0;
        const uint32_t lev = sbits >> 2;
        const uint32_t deg = rend - urow;
        const uint32_t child_bits = ((lev + 1) << 2);
        const bool one_chunk = deg <= 64;
        fw_mask = 0;
        fw_fin = u;
        for (uint32_t base = 0; base < deg && !full; base += 64) {
            const bool act = base + static_cast<uint32_t>(lane) < deg;
            const uint32_t off = urow + base + static_cast<uint32_t>(lane);
            uint32_t cr = cr0, rsc = rs0;
            int32_t wt = w0;
            mem_order();  // every earlier step's state writes are visible to the loads below
            if (base > 0) {
                const uint32_t o = act ? off : urow;
                cr = a.colr[o];
                wt = a.w[o];
                rsc = a.rs[o];
            }
            const uint32_t c = act ? (cr & ~kRootBit) : u;
            const uint4 sc = st[c];
            if (base == 0 && one_chunk && nu != kNone) {
                const uint4 p4 = st[nu];
                const uint32_t pbr = ax[nu].x;
                const uint32_t pre = a.ro[nu + 1];
                const uint32_t poff = min(nrow + static_cast<uint32_t>(lane), a.m2 - 1);
                pf_cr = a.colr[poff];
                pf_w = a.w[poff];
                pf_rs = a.rs[poff];
                pf_bits = uni(p4.w);
                pf_br = uni(pbr);
                pf_rend = uni(pre);
                pf_v = nu;
                pf_ok = true;
            }
            if (base == 0 && lane == 0) {  // finalise u (after the loads above: none of them reads it)
                reinterpret_cast<uint32_t *>(&st[u])[3] = sbits | kIn;
                ax[u].y = nfin;
                if (!bu) ol[nopen] = u;
            }
            const bool reached = act && sc.x == epoch;
            const bool live = act && c != u && !(reached && (sc.w & kIn));
            const int32_t cand = du + wt;
            const bool nb = bu || ((cr & kRootBit) && c < src);
            const bool tb0 = reached && (sc.w & kBlk);
            const uint32_t cprev = __shfl_up(c, 1);
            const bool dup = __ballot(act && lane > 0 && c == cprev) != 0;
            uint64_t improve, writer;
            if (!dup) {
                const bool imp = live && (!reached || cand < static_cast<int32_t>(sc.y));
                improve = __ballot(imp);
                writer = improve;
                open += static_cast<int32_t>(__popcll(__ballot(imp && !nb))) -
                        static_cast<int32_t>(__popcll(__ballot(imp && reached && !tb0)));
            } else {
                // parallel edges in this chunk: the sequential relaxation, lane by lane
                improve = writer = 0;
                uint32_t run_c = kNone;
                bool run_reached = false, run_tb = false;
                int32_t run_d = 0;
                int last = -1;
                const uint64_t livem = __ballot(live);
                const uint64_t reachm = __ballot(reached), tbm = __ballot(tb0), nbm = __ballot(nb);
                const int cntl = static_cast<int>(min(64u, deg - base));
                for (int l = 0; l < cntl; ++l) {
                    if (!((livem >> l) & 1ull)) continue;
                    const uint32_t cl = lane_u32(c, l);
                    if (cl != run_c) {
                        run_c = cl;
                        run_reached = (reachm >> l) & 1ull;
                        run_d = static_cast<int32_t>(lane_u32(sc.y, l));
                        run_tb = (tbm >> l) & 1ull;
                        last = -1;
                    }
                    const int32_t cd = static_cast<int32_t>(lane_u32(static_cast<uint32_t>(cand), l));
                    if (!run_reached || cd < run_d) {
                        const bool nbl = (nbm >> l) & 1ull;
                        open += run_reached ? (static_cast<int32_t>(!nbl) - static_cast<int32_t>(!run_tb))
                                            : static_cast<int32_t>(!nbl);
                        run_reached = true;
                        run_tb = nbl;
                        run_d = cd;
                        improve |= 1ull << l;
                        if (last >= 0) writer &= ~(1ull << last);
                        writer |= 1ull << l;
                        last = l;
                    }
                }
            }
            const uint32_t wbits = child_bits | (nb ? kBlk : 0u);
            const uint32_t wbr = (u == src) ? c : bru;
            if ((writer >> lane) & 1ull) {
                st[c] = make_uint4(epoch, static_cast<uint32_t>(cand), off, wbits);
                ax[c].x = wbr;
            }
            fw_mask = writer;
            fw_c = c;
            fw_bits = wbits;
            fw_br = wbr;
            while (improve) {
                const int l = __ffsll(static_cast<unsigned long long>(improve)) - 1;
                improve &= improve - 1;
                if (!heap_push(hp, hr, len, H, static_cast<int32_t>(lane_u32(static_cast<uint32_t>(cand), l)),
                               lane_u32(c, l), lane_u32(rsc, l), td, tv, tr, lane)) {
                    full = true;
                    break;
                }
            }
        }
        ++nfin;
        if (!bu) {
            --open;
            ++nopen;
        }
        if (full) return 1;
        if (open == 0) break;
    }
    mem_order();
    // accepted candidates: rows of the finalised unblocked vertices, edges to a larger vertex (a loop once)
    if (lane == 0) {
        cnt[0] = 0;
        cnt[1] = 0;
    }
    lds_order();
    for (uint32_t i0 = 0; i0 < nopen; i0 += 64) {
        const uint32_t i = i0 + static_cast<uint32_t>(lane);
        if (i < nopen) {
            const uint32_t x = ol[i];
            const uint4 sx = st[x];
            const uint32_t brx = ax[x].x;
            const uint32_t e1 = a.ro[x + 1];
            for (uint32_t e = a.ro[x]; e < e1; ++e) {
                const uint32_t y = a.colr[e] & ~kRootBit;
                if (y < x) continue;
                const uint32_t re = a.rev[e];
                if (y == x && e > re) continue;
                const uint4 sy = st[y];
                if (sy.x != epoch || !(sy.w & kIn) || (sy.w & kBlk)) continue;
                if (sy.z == e || sx.z == re) continue;
                if (!(x == src || y == src || brx != ax[y].x)) continue;
                const uint32_t k = atomicAdd(&cnt[0], 1u);
                cb[k] = make_uint2(e, static_cast<uint32_t>(static_cast<int32_t>(sx.y) + static_cast<int32_t>(sy.y) +
                                                            a.w[e]));
            }
        }
    }
    lds_order();
    const uint32_t nc = uni(cnt[0]);
    if (lane == 0) kp[src] = epoch;
    mem_order();
    // layout vertices: candidate endpoints and their ancestors, each listed by the lane that marks it
    for (uint32_t i0 = 0; i0 < 2 * nc; i0 += 64) {
        const uint32_t i = i0 + static_cast<uint32_t>(lane);
        if (i < 2 * nc) {
            const uint32_t e = cb[i >> 1].x;
            uint32_t x = (i & 1u) ? (a.colr[e] & ~kRootBit) : a.rows[e];
            while (x != src) {
                if (atomicExch(&kp[x], epoch) == epoch) break;
                const uint4 sx = st[x];
                const uint32_t k = atomicAdd(&cnt[1], 1u);
                kb[k] = make_uint4(x, sx.z, sx.w >> 2, ax[x].y);
                if (sx.z >= a.m2) break;  // cannot happen for a finalised vertex other than the root
                x = a.rows[sx.z];
            }
        }
    }
    mem_order();
    lds_order();
    const uint32_t nk = uni(cnt[1]);
    unsigned long long c_off = 0, k_off = 0;
    if (lane == 0) {
        c_off = atomicAdd(&a.pool_top[0], static_cast<unsigned long long>(nc));
        k_off = atomicAdd(&a.pool_top[1], static_cast<unsigned long long>(nk));
    }
    c_off = uni64(c_off);
    k_off = uni64(k_off);
    const bool fits = c_off + nc <= a.cand_cap && k_off + nk <= a.kept_cap;
    if (fits) {
        for (uint32_t i = static_cast<uint32_t>(lane); i < nc; i += 64) {
            const uint2 cv = cb[i];
            a.cand_edge_pool[c_off + i] = cv.x;
            a.cand_weight_pool[c_off + i] = static_cast<int32_t>(cv.y);
        }
        for (uint32_t i = static_cast<uint32_t>(lane); i < nk; i += 64) a.kept_pool[k_off + i] = kb[i];
    }
    if (lane == 0) {
        a.cand_start[r] = static_cast<int64_t>(c_off);
        a.cand_count[r] = static_cast<int32_t>(nc);
        a.kept_start[r] = static_cast<int64_t>(k_off);
        a.kept_count[r] = static_cast<int32_t>(nk);
    }
    return fits ? 0 : 2;
}

__global__ void __launch_bounds__(64) k_sp_trees(SpArgs a) {
    extern __shared__ __align__(16) uint32_t sp_lds[];
    __shared__ uint32_t cnt[2];
    const uint32_t H = a.heap_cap;
    uint2 *hp = reinterpret_cast<uint2 *>(sp_lds);
    uint32_t *hr = sp_lds + 2 * H;
    const int lane = static_cast<int>(threadIdx.x);
    const size_t slab = blockIdx.x;
    uint4 *st = a.state + slab * a.n;
    uint2 *ax = a.aux + slab * a.n;
    uint32_t *kp = a.keep + slab * a.n;
    uint32_t *ol = a.open_list + slab * a.n;
    uint2 *cb = a.cand_buf + slab * a.cand_stride;
    uint4 *kb = a.kept_buf + slab * a.n;
    for (;;) {
        uint32_t wi = 0;
        if (lane == 0) wi = atomicAdd(a.next, 1u);
        wi = uni(__shfl(wi, 0));
        if (wi >= a.num_work) break;
        const uint32_t r = uni(a.work[wi]);
        const int s = search_root(a, r, hp, hr, cnt, st, ax, kp, ol, cb, kb, lane);
        if (lane == 0) a.status[r] = s;
        mem_order();
    }
}

template <typename T>
hipError_t dalloc(T **p, size_t count) {
    return hipMalloc(reinterpret_cast<void **>(p), std::max<size_t>(count, 1) * sizeof(T));
}

}  // namespace

struct mcb_sptree {
    int device = 0;
    int cus = 0;
    uint32_t n = 0, m2 = 0, F = 0;
    uint32_t *ro = nullptr, *colr = nullptr, *rs = nullptr, *rows = nullptr, *rev = nullptr, *roots = nullptr;
    int32_t *w = nullptr;
    uint32_t *work = nullptr, *next = nullptr;
    uint4 *state = nullptr;
    uint2 *aux = nullptr;
    uint32_t *keep = nullptr, *open_list = nullptr;
    uint2 *cand_buf = nullptr;
    uint4 *kept_buf = nullptr;
    uint64_t slabs = 0, cand_stride = 0;
    int32_t *status = nullptr, *cand_count = nullptr, *kept_count = nullptr;
    int64_t *cand_start = nullptr, *kept_start = nullptr;
    unsigned long long *pool_top = nullptr;
    uint32_t *cand_edge_pool = nullptr;
    int32_t *cand_weight_pool = nullptr;
    uint4 *kept_pool = nullptr;
    uint64_t cand_cap = 0, kept_cap = 0;
    uint32_t epoch_base = 0;
    int waves_per_cu[kStages] = {8, 2, 1};
    uint32_t lds_per_cu = 0, lds_max = 0, lds_static = 0;  // bytes: a CU's LDS, a workgroup's limit, static use
    hipStream_t stream = nullptr;
    // results on the host
    std::vector<int32_t> h_status, h_cand_count, h_kept_count;
    std::vector<int64_t> h_cand_start, h_kept_start;
    std::vector<uint32_t> h_cand_edge;
    std::vector<int32_t> h_cand_weight;
    std::vector<uint32_t> h_kept;
    mcb_sptree_stats stats{};
    bool have_results = false;

    ~mcb_sptree() {
        for (void *p : {static_cast<void *>(ro), static_cast<void *>(colr), static_cast<void *>(rs),
                        static_cast<void *>(rows), static_cast<void *>(rev), static_cast<void *>(roots),
                        static_cast<void *>(w), static_cast<void *>(work), static_cast<void *>(next),
                        static_cast<void *>(state), static_cast<void *>(aux), static_cast<void *>(keep),
                        static_cast<void *>(open_list), static_cast<void *>(cand_buf), static_cast<void *>(kept_buf),
                        static_cast<void *>(status), static_cast<void *>(cand_count), static_cast<void *>(kept_count),
                        static_cast<void *>(cand_start), static_cast<void *>(kept_start), static_cast<void *>(pool_top),
                        static_cast<void *>(cand_edge_pool), static_cast<void *>(cand_weight_pool),
                        static_cast<void *>(kept_pool)})
            if (p) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

// LDS heap entries of a stage: the workgroup's share of the CU's LDS less the kernel's static LDS,
// 12 bytes per entry.
uint32_t stage_heap_cap(const mcb_sptree *t, int stage) {
    const uint32_t share = t->lds_per_cu / static_cast<uint32_t>(t->waves_per_cu[stage]);
    const uint32_t bytes = std::min(share, t->lds_max) - t->lds_static;
    return bytes / 12u;
}

// Grows a pool to hold twice the reserved count: the roots searched again reserve at most what they
// reserved before, which is part of that count, so one retry always fits.
int grow_pool(mcb_sptree *t, int which, uint64_t need) {
    if (which == 0) {
        const uint64_t cap = 2 * need + 1024;
        uint32_t *e = nullptr;
        int32_t *w = nullptr;
        SP_TRY(dalloc(&e, cap));
        SP_TRY(dalloc(&w, cap));
        SP_TRY(hipMemcpyAsync(e, t->cand_edge_pool, t->cand_cap * sizeof(uint32_t), hipMemcpyDeviceToDevice,
                              t->stream));
        SP_TRY(hipMemcpyAsync(w, t->cand_weight_pool, t->cand_cap * sizeof(int32_t), hipMemcpyDeviceToDevice,
                              t->stream));
        SP_TRY(hipStreamSynchronize(t->stream));
        SP_TRY(hipFree(t->cand_edge_pool));
        SP_TRY(hipFree(t->cand_weight_pool));
        t->cand_edge_pool = e;
        t->cand_weight_pool = w;
        t->cand_cap = cap;
    } else {
        const uint64_t cap = 2 * need + 1024;
        uint4 *k = nullptr;
        SP_TRY(dalloc(&k, cap));
        SP_TRY(hipMemcpyAsync(k, t->kept_pool, t->kept_cap * sizeof(uint4), hipMemcpyDeviceToDevice, t->stream));
        SP_TRY(hipStreamSynchronize(t->stream));
        SP_TRY(hipFree(t->kept_pool));
        t->kept_pool = k;
        t->kept_cap = cap;
    }
    return MCB_OK;
}

}  // namespace

extern "C" {

const char *mcb_sptree_last_error(void) { return g_sp_err.c_str(); }

int mcb_sptree_create(int32_t nodes, int32_t directed_edges, const uint32_t *row_offsets, const uint32_t *rows,
                      const uint32_t *cols, const int32_t *weights, const uint32_t *reverse, const int32_t *fvs_rank,
                      const uint32_t *roots, int32_t num_roots, mcb_sptree **out) {
    if (!out || nodes <= 0 || directed_edges <= 0 || num_roots <= 0 || !row_offsets || !rows || !cols || !weights ||
        !reverse || !fvs_rank || !roots)
        return sfail(MCB_E_INVALID, "invalid arguments");
    if (static_cast<uint32_t>(nodes) >= kRootBit) return sfail(MCB_E_INVALID, "too many vertices");
    *out = nullptr;
    const uint32_t n = static_cast<uint32_t>(nodes), m2 = static_cast<uint32_t>(directed_edges);
    if (row_offsets[n] != m2) return sfail(MCB_E_INVALID, "row offsets do not end at the edge count");
    for (uint32_t e = 0; e < m2; ++e)
        if (rows[e] >= n || cols[e] >= n || reverse[e] >= m2) return sfail(MCB_E_INVALID, "edge endpoint out of range");
    for (int32_t r = 0; r < num_roots; ++r)
        if (roots[r] >= n) return sfail(MCB_E_INVALID, "root out of range");
    std::unique_ptr<mcb_sptree> t(new (std::nothrow) mcb_sptree());
    if (!t) return sfail(MCB_E_NOMEM, "out of memory");
    SP_TRY(hipGetDevice(&t->device));
    hipDeviceProp_t prop;
    SP_TRY(hipGetDeviceProperties(&prop, t->device));
    t->cus = prop.multiProcessorCount;
    hipFuncAttributes fattr{};
    SP_TRY(hipFuncGetAttributes(&fattr, reinterpret_cast<const void *>(&k_sp_trees)));
    t->lds_static = static_cast<uint32_t>(fattr.sharedSizeBytes) + 64u;
    t->lds_per_cu = 160u * 1024u;
    t->lds_max = static_cast<uint32_t>(std::min<size_t>(prop.sharedMemPerBlock, t->lds_per_cu));
    if (t->lds_max <= t->lds_static + 12u * 64u) return sfail(MCB_E_DEVICE, "too little LDS per workgroup");
    SP_TRY(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_sp_trees),
                               hipFuncAttributeMaxDynamicSharedMemorySize,
                               static_cast<int>(t->lds_max - static_cast<uint32_t>(fattr.sharedSizeBytes))));
    t->n = n;
    t->m2 = m2;
    t->F = static_cast<uint32_t>(num_roots);
    std::vector<uint32_t> colr(m2), rs(m2);
    for (uint32_t e = 0; e < m2; ++e) {
        const uint32_t c = cols[e];
        colr[e] = c | (fvs_rank[c] >= 0 ? kRootBit : 0u);
        rs[e] = row_offsets[c];
    }
    SP_TRY(hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking));
    SP_TRY(dalloc(&t->ro, n + 1));
    SP_TRY(dalloc(&t->colr, m2));
    SP_TRY(dalloc(&t->rs, m2));
    SP_TRY(dalloc(&t->rows, m2));
    SP_TRY(dalloc(&t->rev, m2));
    SP_TRY(dalloc(&t->w, m2));
    SP_TRY(dalloc(&t->roots, t->F));
    SP_TRY(dalloc(&t->work, t->F));
    SP_TRY(dalloc(&t->next, 1));
    SP_TRY(hipMemcpy(t->ro, row_offsets, (n + 1) * sizeof(uint32_t), hipMemcpyHostToDevice));
    SP_TRY(hipMemcpy(t->colr, colr.data(), m2 * sizeof(uint32_t), hipMemcpyHostToDevice));
    SP_TRY(hipMemcpy(t->rs, rs.data(), m2 * sizeof(uint32_t), hipMemcpyHostToDevice));
    SP_TRY(hipMemcpy(t->rows, rows, m2 * sizeof(uint32_t), hipMemcpyHostToDevice));
    SP_TRY(hipMemcpy(t->rev, reverse, m2 * sizeof(uint32_t), hipMemcpyHostToDevice));
    SP_TRY(hipMemcpy(t->w, weights, m2 * sizeof(int32_t), hipMemcpyHostToDevice));
    SP_TRY(hipMemcpy(t->roots, roots, t->F * sizeof(uint32_t), hipMemcpyHostToDevice));
    // slabs: one per wave of the widest stage, within a quarter of the free device memory
    t->cand_stride = m2 / 2 + 1;
    const uint64_t per_slab = static_cast<uint64_t>(n) * (sizeof(uint4) + sizeof(uint2) + 2 * sizeof(uint32_t) +
                                                          sizeof(uint4)) +
                              t->cand_stride * sizeof(uint2);
    size_t free_b = 0, total_b = 0;
    SP_TRY(hipMemGetInfo(&free_b, &total_b));
    uint64_t slabs = std::min<uint64_t>(t->F, static_cast<uint64_t>(kMaxWavesPerCu) * t->cus);
    slabs = std::min<uint64_t>(slabs, std::max<uint64_t>(1, (free_b / 4) / per_slab));
    t->slabs = slabs;
    SP_TRY(dalloc(&t->state, slabs * n));
    SP_TRY(dalloc(&t->aux, slabs * n));
    SP_TRY(dalloc(&t->keep, slabs * n));
    SP_TRY(dalloc(&t->open_list, slabs * n));
    SP_TRY(dalloc(&t->kept_buf, slabs * n));
    SP_TRY(dalloc(&t->cand_buf, slabs * t->cand_stride));
    // cleared on the search's own stream: it is non-blocking, so work on the null stream would not be
    // ordered before its launches
    SP_TRY(hipMemsetAsync(t->state, 0, slabs * n * sizeof(uint4), t->stream));
    SP_TRY(hipMemsetAsync(t->keep, 0, slabs * n * sizeof(uint32_t), t->stream));
    SP_TRY(hipStreamSynchronize(t->stream));
    SP_TRY(dalloc(&t->status, t->F));
    SP_TRY(dalloc(&t->cand_count, t->F));
    SP_TRY(dalloc(&t->kept_count, t->F));
    SP_TRY(dalloc(&t->cand_start, t->F));
    SP_TRY(dalloc(&t->kept_start, t->F));
    SP_TRY(dalloc(&t->pool_top, 2));
    // first pool capacities (they grow on demand); MCB_SPTREE_POOL_MIN lowers the floor of 2^20 entries
    uint64_t floor_cap = 1u << 20;
    if (const char *env = std::getenv("MCB_SPTREE_POOL_MIN")) floor_cap = std::max<uint64_t>(1, std::strtoull(env, nullptr, 10));
    t->cand_cap = std::max<uint64_t>(floor_cap, 64ull * t->F);
    t->kept_cap = std::max<uint64_t>(floor_cap, 128ull * t->F);
    if (floor_cap < (1u << 20)) {
        t->cand_cap = floor_cap;
        t->kept_cap = floor_cap;
    }
    SP_TRY(dalloc(&t->cand_edge_pool, t->cand_cap));
    SP_TRY(dalloc(&t->cand_weight_pool, t->cand_cap));
    SP_TRY(dalloc(&t->kept_pool, t->kept_cap));
    t->stats.slabs = static_cast<int64_t>(slabs);
    *out = t.release();
    return MCB_OK;
}

int mcb_sptree_run(mcb_sptree *t, int stages) {
    if (!t) return sfail(MCB_E_INVALID, "invalid arguments");
    if (stages < 1 || stages > kStages) return sfail(MCB_E_INVALID, "stages must be 1..3");
    SP_TRY(hipSetDevice(t->device));
    t->have_results = false;
    mcb_sptree_stats st{};
    st.slabs = static_cast<int64_t>(t->slabs);
    hipEvent_t e0, e1;
    SP_TRY(hipEventCreate(&e0));
    SP_TRY(hipEventCreate(&e1));
    SP_TRY(hipEventRecord(e0, t->stream));
    // epochs stay distinct across launches; clear the slabs before the 32-bit stamp wraps
    const uint64_t launches_left = (0xFFFFFFFFull - t->epoch_base) / (t->F + 1);
    if (launches_left < 2 * kStages + 2) {
        SP_TRY(hipMemsetAsync(t->state, 0, t->slabs * t->n * sizeof(uint4), t->stream));
        SP_TRY(hipMemsetAsync(t->keep, 0, t->slabs * t->n * sizeof(uint32_t), t->stream));
        t->epoch_base = 0;
    }
    std::vector<uint32_t> work(t->F);
    for (uint32_t r = 0; r < t->F; ++r) work[r] = r;  // ascending root id: the widest searches first
    std::vector<int32_t> status(t->F, 1);
    SP_TRY(hipMemsetAsync(t->pool_top, 0, 2 * sizeof(unsigned long long), t->stream));
    int stage = 0;
    while (!work.empty() && stage < stages) {
        const uint32_t H = stage_heap_cap(t, stage);
        const uint64_t grid =
            std::min<uint64_t>({static_cast<uint64_t>(work.size()), t->slabs,
                                static_cast<uint64_t>(t->waves_per_cu[stage]) * t->cus});
        SP_TRY(hipMemcpyAsync(t->work, work.data(), work.size() * sizeof(uint32_t), hipMemcpyHostToDevice, t->stream));
        SP_TRY(hipMemsetAsync(t->next, 0, sizeof(uint32_t), t->stream));
        SpArgs a{};
        a.ro = t->ro;
        a.colr = t->colr;
        a.w = t->w;
        a.rs = t->rs;
        a.rows = t->rows;
        a.rev = t->rev;
        a.roots = t->roots;
        a.work = t->work;
        a.n = t->n;
        a.m2 = t->m2;
        a.num_work = static_cast<uint32_t>(work.size());
        a.heap_cap = H;
        a.epoch_base = t->epoch_base;
        a.next = t->next;
        a.state = t->state;
        a.aux = t->aux;
        a.keep = t->keep;
        a.open_list = t->open_list;
        a.cand_buf = t->cand_buf;
        a.kept_buf = t->kept_buf;
        a.cand_stride = t->cand_stride;
        a.status = t->status;
        a.cand_start = t->cand_start;
        a.cand_count = t->cand_count;
        a.kept_start = t->kept_start;
        a.kept_count = t->kept_count;
        a.pool_top = t->pool_top;
        a.cand_edge_pool = t->cand_edge_pool;
        a.cand_weight_pool = t->cand_weight_pool;
        a.kept_pool = t->kept_pool;
        a.cand_cap = t->cand_cap;
        a.kept_cap = t->kept_cap;
        hipLaunchKernelGGL(k_sp_trees, dim3(static_cast<uint32_t>(grid)), dim3(64), H * 12u, t->stream, a);
        SP_TRY(hipGetLastError());
        t->epoch_base += t->F + 1;
        st.launches++;
        std::vector<int32_t> s_dev(t->F);
        unsigned long long top[2];
        SP_TRY(hipMemcpyAsync(s_dev.data(), t->status, t->F * sizeof(int32_t), hipMemcpyDeviceToHost, t->stream));
        SP_TRY(hipMemcpyAsync(top, t->pool_top, sizeof(top), hipMemcpyDeviceToHost, t->stream));
        SP_TRY(hipStreamSynchronize(t->stream));
        std::vector<uint32_t> again, next_stage;
        for (uint32_t wi : work) {
            status[wi] = s_dev[wi];
            if (s_dev[wi] == 2) again.push_back(wi);
            else if (s_dev[wi] == 1) next_stage.push_back(wi);
        }
        st.stage_roots[stage] += static_cast<int64_t>(work.size() - again.size() - next_stage.size());
        if (!again.empty()) {  // pools full: grow them (keeping what is there) and search those roots again
            if (top[0] > t->cand_cap) {
                int rc = grow_pool(t, 0, top[0]);
                if (rc) return rc;
            }
            if (top[1] > t->kept_cap) {
                int rc = grow_pool(t, 1, top[1]);
                if (rc) return rc;
            }
            // the reservations of the failed roots stay as holes; the counters keep counting from the top
            work = again;
            continue;
        }
        // a stage that finished under an eighth of its roots says the rest need heaps far beyond LDS
        // (an expander, where every root reaches the whole graph): leave them to the host at once
        const size_t finished = work.size() - next_stage.size();
        work = next_stage;
        ++stage;
        if (finished * 8 < finished + work.size()) break;
    }
    st.host_roots = static_cast<int64_t>(work.size());
    SP_TRY(hipEventRecord(e1, t->stream));
    SP_TRY(hipEventSynchronize(e1));
    float ms = 0;
    SP_TRY(hipEventElapsedTime(&ms, e0, e1));
    st.device_ms = ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    // results to the host
    unsigned long long top[2];
    SP_TRY(hipMemcpy(top, t->pool_top, sizeof(top), hipMemcpyDeviceToHost));
    const uint64_t nc = std::min<uint64_t>(top[0], t->cand_cap), nk = std::min<uint64_t>(top[1], t->kept_cap);
    try {
        t->h_status = status;
        t->h_cand_start.resize(t->F);
        t->h_cand_count.resize(t->F);
        t->h_kept_start.resize(t->F);
        t->h_kept_count.resize(t->F);
        t->h_cand_edge.resize(nc);
        t->h_cand_weight.resize(nc);
        t->h_kept.resize(4 * nk);
    } catch (const std::bad_alloc &) {
        return sfail(MCB_E_NOMEM, "out of host memory for the search results");
    }
    SP_TRY(hipMemcpy(t->h_cand_start.data(), t->cand_start, t->F * sizeof(int64_t), hipMemcpyDeviceToHost));
    SP_TRY(hipMemcpy(t->h_cand_count.data(), t->cand_count, t->F * sizeof(int32_t), hipMemcpyDeviceToHost));
    SP_TRY(hipMemcpy(t->h_kept_start.data(), t->kept_start, t->F * sizeof(int64_t), hipMemcpyDeviceToHost));
    SP_TRY(hipMemcpy(t->h_kept_count.data(), t->kept_count, t->F * sizeof(int32_t), hipMemcpyDeviceToHost));
    SP_TRY(hipMemcpy(t->h_cand_edge.data(), t->cand_edge_pool, nc * sizeof(uint32_t), hipMemcpyDeviceToHost));
    SP_TRY(hipMemcpy(t->h_cand_weight.data(), t->cand_weight_pool, nc * sizeof(int32_t), hipMemcpyDeviceToHost));
    SP_TRY(hipMemcpy(t->h_kept.data(), t->kept_pool, nk * sizeof(uint4), hipMemcpyDeviceToHost));
    for (uint32_t r = 0; r < t->F; ++r)
        if (status[r] != 0) {
            t->h_cand_start[r] = t->h_kept_start[r] = 0;
            t->h_cand_count[r] = t->h_kept_count[r] = 0;
        }
    st.candidates = static_cast<int64_t>(nc);
    st.kept = static_cast<int64_t>(nk);
    t->stats = st;
    t->have_results = true;
    return MCB_OK;
}

int mcb_sptree_set_stages(mcb_sptree *t, const int32_t *waves_per_cu, int count) {
    if (!t || !waves_per_cu || count < 1 || count > kStages) return sfail(MCB_E_INVALID, "invalid arguments");
    for (int i = 0; i < count; ++i) {
        const int wpc = waves_per_cu[i];
        if (wpc < 1 || wpc > kMaxWavesPerCu || (i > 0 && wpc > waves_per_cu[i - 1]))
            return sfail(MCB_E_INVALID, "waves per CU must be 1..16 and not increase from stage to stage");
        if (static_cast<uint64_t>(wpc) * t->cus > t->slabs && static_cast<uint64_t>(wpc) * t->cus < t->F)
            return sfail(MCB_E_INVALID, "more waves than the workspaces made at create");
        if (t->lds_per_cu / static_cast<uint32_t>(wpc) <= t->lds_static + 12u * 64u)
            return sfail(MCB_E_INVALID, "too many waves per CU for a heap of 64 entries");
    }
    for (int i = 0; i < kStages; ++i) t->waves_per_cu[i] = waves_per_cu[std::min(i, count - 1)];
    return MCB_OK;
}

int mcb_sptree_results(const mcb_sptree *t, mcb_root_trees *view) {
    if (!t || !view) return sfail(MCB_E_INVALID, "invalid arguments");
    if (!t->have_results) return sfail(MCB_E_STATE, "no results: call mcb_sptree_run first");
    view->status = t->h_status.data();
    view->cand_start = t->h_cand_start.data();
    view->cand_count = t->h_cand_count.data();
    view->cand_edge = t->h_cand_edge.data();
    view->cand_weight = t->h_cand_weight.data();
    view->kept_start = t->h_kept_start.data();
    view->kept_count = t->h_kept_count.data();
    view->kept = t->h_kept.data();
    return MCB_OK;
}

int mcb_sptree_get_stats(const mcb_sptree *t, mcb_sptree_stats *stats) {
    if (!t || !stats) return sfail(MCB_E_INVALID, "invalid arguments");
    *stats = t->stats;
    return MCB_OK;
}

void mcb_sptree_destroy(mcb_sptree *t) { delete t; }

}  // extern "C"
