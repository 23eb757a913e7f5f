// dense.hip -- exact reduced row echelon form of a dense block mod p.
//
// Replaces exact_dense_linear_algebra_ff_{8,16,32} (la_ff_32.c:1217-1333),
// interreduce_dense_matrix_ff_{8,16,32} (la_ff_32.c:1181-1215) and their
// row kernels reduce_dense_row_by_dense_new_pivots_{17,31}_bit
// (la_ff_32.c:632-753).  The reference reduces rows one at a time against
// published pivots and back-substitutes at the end; the result is the
// unique reduced echelon form of the row space, so any elimination order
// gives the same bits.  This file uses blocked Gauss-Jordan elimination,
// which turns almost all of the work into one product per 64-column panel:
//
//   for each panel of 64 columns [c0, c0 + 64):
//     select  (1 workgroup): scan the non-pivot rows in order, keep the
//             panel parts of the rows that add rank, in reduced form E, and
//             the transform T with E = T * A[sel, panel];  k <= 64 pivots
//     rp      Rp = T * A[sel, c0:n]   (rows normalised, identity on the
//             panel's pivot columns); written back into the selected rows
//     gather  L[r] = A[r, pivot columns] for every other row
//     update  A[r, c0:n] -= L[r] * Rp  for every other row with L[r] != 0
//             -- previous pivot rows too, so the result is fully reduced
//
// Pivot rows stay where they are; rowpiv[r] / colpiv[c] record the pairing.
// The update is the hot kernel: plain form on the integer ALU (64-bit
// multiply-add), MFMA form on the FP64 matrix cores with the exactness
// bound k (p-1)^2 < 2^53 (p < 2^17) or, for larger p, 16-bit operand
// splitting (four exact products recombined in integers).
#include <chrono>
#include <climits>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>

#include "internal.h"

namespace gbgpu {

constexpr int PB = 64;  // panel width

struct Panel {
    int32_t k;           // pivots found in this panel
    int32_t erow[PB];    // selected rows
    int32_t ecol[PB];    // their pivot columns (absolute)
    uint32_t T[PB * PB]; // Rp[t] = sum_s T[t][s] * A[erow[s]]
};

// the candidate rows: cand[0, head) are pivot rows, cand[head, count)
// candidates (in any order)
struct ListState {
    int32_t done;   // set once no candidate row is left
    int32_t head;
    int32_t count;
    int32_t nzero;  // candidates found zero on their panel since the last prune
    int32_t pcount; // blocks of a prune that have written their flags
};

// Ordering of the panel loop's two streams through device words (the MFMA
// form).  A marker and a wait between the streams cost several
// microseconds of dispatch each, twice a panel, on the loop's critical
// path; instead a kernel posts a word and one thread of the kernel that
// depends on it waits for it:
//   w[0]    the k_panel_narrow launch of panel j has begun, so the select of
//           panel j before it on the main stream is complete (s0 + j);
//   w[32]   the select of panel j has begun, so the k_panel_narrow launch
//           of panel j - 1 before it is complete (s0 + j);
//   w[64]   the side stream's work of panel j is complete (s0 + j);
//   w[96]   a wait ran out of its bound: the s0 of that loop.
// Every wait is for a kernel queued by the host before the waiting one: if
// the two streams share a hardware queue, which runs them in that order,
// what a wait needs has run by then.  A tool that runs kernels one at a
// time in another order (rocprofv3 --pmc does) makes the waits run out:
// such runs take GBGPU_DENSE_SYNC=events.
// The side stream's update waits (through its Rp launch) for the next
// select to have begun, so that the select holds its CU before the
// update's blocks fill the device.  Only single blocks wait -- one-block
// launches on either stream and one block of the side stream's Rp launch
// -- and what each waits for never
// waits on the side stream's later work, nor for the device to have room
// for more than one waiting block, so the waits end; each is still
// bounded, and running out of the bound marks the loop failed (herr) and
// makes every later wait of the loop return at once.  Releases and
// acquires are at agent scope: the words and the data they publish cross
// the XCDs' L2s.
struct LoopSync {
    int32_t *w = nullptr;      // nullptr: the loop orders its streams by events
    int32_t *herr = nullptr;   // mapped host word, 1 once a wait ran out
    int32_t *hprog = nullptr;  // mapped host word, the last w[64] posted
    int32_t s0 = 0;
};
constexpr int LS_NAR = 0, LS_SEL = 32, LS_SIDE = 64, LS_ABORT = 96;
constexpr uint32_t LS_SPIN = 1u << 22;  // polls of a wait (about a microsecond each)

__device__ __forceinline__ bool seq_reached(int32_t v, int32_t want) {
    return (int32_t)((uint32_t)v - (uint32_t)want) >= 0;
}

__device__ __forceinline__ void ls_post(const LoopSync &ls, int at, int32_t v) {
    __hip_atomic_store(ls.w + at, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// one thread: until w[at] reaches want (polled by coherent loads that leave
// the caches alone, one acquire once it has)
__device__ void ls_wait(const LoopSync &ls, int at, int32_t want) {
    for (uint32_t i = 0;; ++i) {
        if (seq_reached(__hip_atomic_load(ls.w + at, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), want)) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            return;
        }
        if (seq_reached(__hip_atomic_load(ls.w + LS_ABORT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), ls.s0)) return;
        if (i >= LS_SPIN) {
            __hip_atomic_store(ls.w + LS_ABORT, ls.s0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (ls.herr) __hip_atomic_store(ls.herr, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            return;
        }
        __builtin_amdgcn_s_sleep(2);
    }
}

__device__ uint32_t inv_mod(uint32_t a, uint32_t p) {
    // extended Euclid on 32-bit operands (tools.c:83-110 computes the same
    // inverse); |s| stays below p < 2^31
    uint32_t r0 = p, r1 = a % p;
    int32_t s0 = 0, s1 = 1;
    while (r1 != 0) {
        const uint32_t q = r0 / r1;
        const uint32_t t = r0 - q * r1;
        r0 = r1;
        r1 = t;
        const int32_t u = (int32_t)((int64_t)s0 - (int64_t)q * s1);
        s0 = s1;
        s1 = u;
    }
    return (uint32_t)(s0 < 0 ? s0 + (int64_t)p : s0);
}

__device__ __forceinline__ uint32_t sub_mod(uint32_t a, uint32_t b, uint32_t p) {
    return a >= b ? a - b : a + p - b;
}

// ------------------------------------------------------------ select
//
// One workgroup finds the panel's pivots.  An LDS array of PB slots holds
// rows of [64 panel entries | transform over the slots]: the pivot rows
// found so far (E) and a tile of candidate rows (rows that are no pivot yet,
// kept in the list cand[head, count) in any order) loaded into the free
// slots in list order, each with a unit transform.  The tile is first
// reduced by E (normalised and reduced within the panel).  Then the panel's
// columns are swept from the left in sub-panels of SW columns:
//
//   chain (one wave, lane = slot, the sub-panel's entries in registers):
//     for each column c of the sub-panel, the lowest unselected tile row
//     nonzero there is the pivot (columns left of c are zero in every
//     unselected row), with pivot value a; every other slot x becomes
//     a x - x[c] * pivot row, without division: no inverse sits on the
//     per-pivot path.  Each slot's new row is kept as a combination of the
//     rows at the start of the sub-panel: D[s] * row s + sum_c C[s][c] *
//     (row of the pivot of column c);
//   apply (every wave, lane = slot, a wave per column): the same
//     combination on every column, the sub-panel's and the transform part
//     included -- a 64 x 64 by 64 x SW product per sub-panel instead of a
//     barrier per pivot.  (The chain only keeps the entries right of the
//     current column: every other row is scaled by the pivot value, which
//     the combination carries.)
//
// At the end of a tile the E rows are normalised, their inverses computed
// in parallel.  Tile rows that are never selected end up zero on the panel.
// A row's transform only ever takes in pivot rows, so the transform T of E
// over the selected rows gives E = T * A[selected rows, panel], and Rp =
// T * A[selected rows, c0:n] is in reduced form with an identity on the
// panel's pivot columns.  Selected rows are swapped to the front of the
// candidate list.

constexpr int XW = 2 * PB + 1;  // padded slot row: panel part, transform part

// a x + b y mod p for p < 2^16, residues a, x, y and b <= p: both products
// are below 2^32 (24-bit multiplies), the quotient of their sum by p is
// estimated in single precision with an absolute error below 0.05, so the
// remainder r computed in wrapped 32-bit arithmetic lies in [-p, 2p); of r,
// r + p and r - p as unsigned words the smallest is the residue
__device__ __forceinline__ uint32_t mad2_mod16(uint32_t a, uint32_t x, uint32_t b, uint32_t y,
                                               uint32_t p, float pinv) {
    const uint32_t s = __umul24(a, x), t = __umul24(b, y);
    const uint32_t q = (uint32_t)(((float)s + (float)t) * pinv);
    const uint32_t r = s + t - __umul24(q, p);
    return min(min(r, r + p), r - p);
}

// a x + b y mod p for residues a, x, y and b <= p (two products below 2^62)
template <int P16>
__device__ __forceinline__ uint32_t sel_mad(uint32_t a, uint32_t x, uint32_t b, uint32_t y,
                                            const Mod &md, float pinv) {
    if (P16 == 1) return mad2_mod16(a, x, b, y, md.p, pinv);
    return md.reduce((uint64_t)a * x + (uint64_t)b * y);
}

constexpr int SEL_T = 512;            // threads of the select workgroup
constexpr int SEL_W = SEL_T / 64;     // its waves
constexpr int SW = 32;                // columns of a sub-panel of the pivot chain

struct SelShared {
    uint32_t X[PB][XW];
    uint32_t C[PB][PB + 1];           // combination coefficients of each slot
    uint32_t D[PB];                   // ... and its own scale
    int32_t qs[PB];                   // the slots the coefficients refer to
    uint32_t einv[PB];
    int32_t ecol[PB], eslot[PB], erow[PB], epos[PB], srow[PB], spos[PB];
    // the tile's reduction by E: coefficients and slots (kept apart from the
    // chain's, the transform part is reduced after a chain)
    uint32_t CE[PB][PB + 1];
    int32_t qse[PB];
    int32_t nzw[SEL_W];               // waves with a nonzero on the sub-panel
    // each chain step's column entries (one per slot), tagged with the
    // chain and step in the high word: a wave polls its entries until every
    // lane shows the step's tag instead of meeting the others at a barrier
    uint64_t f[SW][PB];
    // the move of the selected rows to the front of the candidate list
    int32_t fr[PB], fpos[PB], mvj[PB], mvv[PB];
    // signed base-256 digit planes of a sub-panel's combination (slot x
    // term) and of the pivot rows it takes in (column x term), for the
    // products on the matrix cores
    int8_t Cd[4][PB][SW];
    int8_t Pd[4][2 * PB][SW];
};

// X[s][col] <- D[s] X[s][col] + sum_{q < nq} Cm[s][q] X[qm[q]][col] for every
// slot s, over the panel columns in the mask pcm and, with tr, the
// transform part (D[s] = 1 with unitd).  Lane = slot; a wave takes every
// SEL_W-th column: a column's entries of every slot are read by one wave
// before it writes them, and no other wave touches that column.  The terms
// go in chunks of 16 with no branch inside a chunk (Cm is zero past nq up to
// the chunk's end), so a chunk's 16 reads are in flight together.  KIND 0
// (p < 2^16): products below 2^32 from 24-bit multiplies, summed in 64
// bits; KIND 1 (p < 2^17): products below 2^34, 65 of them fit in 64 bits;
// KIND 2: sums carry into hi.
constexpr int QC = 16;  // terms per chunk

template <int NQ, int KIND>
__device__ __forceinline__ void sel_apply(SelShared &sh, const uint32_t (*Cm)[PB + 1],
                                          const int32_t *qm, bool unitd, int nq, uint64_t pcm,
                                          bool tr, const Mod &md, uint32_t r64) {
    static_assert(NQ % QC == 0, "whole chunks");
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    nq = __builtin_amdgcn_readfirstlane(nq);
    // one chunk's coefficients and row offsets (word offset of the row of
    // slot qm[q]) in registers: loaded once for a single chunk, per chunk
    // and column pair otherwise
    uint32_t cq[QC];
    int32_t qo[QC];
    auto chunk = [&](int k) {
#pragma unroll
        for (int j = 0; j < QC; ++j) {
            cq[j] = Cm[lane][k * QC + j];
            qo[j] = qm[k * QC + j] * XW;
        }
    };
    if (NQ == QC) chunk(0);
    const uint32_t dv = unitd ? 1u : sh.D[lane];
    const uint32_t *xf = &sh.X[0][0];
    // two columns at once (cb = -1: none), so that their reads overlap
    auto two = [&](int ca, int cb) {
        const int cs[2] = {ca, cb < 0 ? ca : cb};  // indexed by unrolled loops only
        uint64_t lo[2];
        uint32_t hi[2] = {0, 0};
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            // __umul24 returns an int: the low 32 bits go through uint32_t,
            // not a sign extension
            lo[u] = KIND == 0 ? (uint64_t)(uint32_t)__umul24(dv, sh.X[lane][cs[u]])
                              : (uint64_t)dv * sh.X[lane][cs[u]];
        }
#pragma unroll
        for (int k = 0; k < NQ / QC; ++k) {
            if (k * QC < nq) {
                if (NQ != QC) chunk(k);
                uint32_t pv[2][QC];
#pragma unroll
                for (int j = 0; j < QC; ++j)
#pragma unroll
                    for (int u = 0; u < 2; ++u) pv[u][j] = xf[qo[j] + cs[u]];
#pragma unroll
                for (int j = 0; j < QC; ++j) {
                    const uint32_t c = cq[j];
#pragma unroll
                    for (int u = 0; u < 2; ++u) {
                        if (KIND == 0) {
                            lo[u] += (uint64_t)(uint32_t)__umul24(c, pv[u][j]);
                        } else if (KIND == 1) {
                            lo[u] += (uint64_t)c * pv[u][j];
                        } else {
                            const uint64_t o = lo[u];
                            lo[u] += (uint64_t)c * pv[u][j];
                            hi[u] += lo[u] < o;
                        }
                    }
                }
            }
        }
        uint32_t r[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            r[u] = md.reduce(lo[u]);
            if (KIND == 2 && hi[u]) {
                r[u] += md.mul(hi[u], r64);
                if (r[u] >= md.p) r[u] -= md.p;
            }
        }
        sh.X[lane][ca] = r[0];
        if (cb >= 0) sh.X[lane][cb] = r[1];
    };
    // panel columns: this wave's are the set bits wv, wv + SEL_W, ... of pcm
    uint64_t m = pcm;
    for (int i = 0; i < wv && m; ++i) m &= m - 1;
    while (m) {
        const int ca = __builtin_ctzll(m);
        for (int i = 0; i < SEL_W && m; ++i) m &= m - 1;
        int cb = -1;
        if (m) {
            cb = __builtin_ctzll(m);
            for (int i = 0; i < SEL_W && m; ++i) m &= m - 1;
        }
        two(ca, cb);
    }
    if (tr)
        for (int col = PB + wv; col < 2 * PB; col += 2 * SEL_W) two(col, col + SEL_W);
}

// the first ND signed base-256 digits of x's representative (x itself when
// x <= hi = 127 (1 + 256 + ...), else x - p), digit i in byte i, the bytes
// above zero
template <int ND>
__device__ __forceinline__ uint32_t digits_n(uint32_t x, uint32_t p, uint32_t hi) {
    int32_t v = x <= hi ? (int32_t)x : (int32_t)x - (int32_t)p;
    uint32_t w = 0;
#pragma unroll
    for (int k = 0; k < ND; ++k) {
        const int32_t d = (int32_t)(int8_t)(v & 0xFF);
        w |= (uint32_t)(d & 0xFF) << (8 * k);
        if (k + 1 < ND) v = (v - d) >> 8;
    }
    return w;
}

// the four signed base-256 digits of x's representative (x itself when
// x <= hi = 127 (1 + 256 + ...), else x - p), digit i in byte i
__device__ __forceinline__ uint32_t digits4(uint32_t x, uint32_t p, uint32_t hi) {
    int32_t v = x <= hi ? (int32_t)x : (int32_t)x - (int32_t)p;
    uint32_t w = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int8_t d = (int8_t)(v & 0xFF);
        w |= (uint32_t)(uint8_t)d << (8 * k);
        v = (v - d) >> 8;
    }
    return w;
}

// The sub-panel apply on the int8 matrix cores: X <- diag(D) X + C P with
// P = X[qs[0..SW)] (SW x 2 PB), every column of X (the panel columns past
// bw are zero in every slot).  C and P go into ND signed digits each; the
// products of digit sum t = i + j are v_mfma_i32_16x16x64_i8 with the pairs
// (i, j) in its four 16-term K groups, a pair taking SW / 16 of them (so
// 64 / SW pairs per product), below 4 * SW * 128^2 <= 2^21 in size.  Wave
// w: slots 16 (w % 4) .., columns 64 (w / 4) .. (four 16 x 16 tiles).
// Every output has one writer, which reads its old value only.
template <int ND>
__device__ __forceinline__ void sel_apply_mfma(SelShared &sh, const Mod &md, const Digits &dg) {
    static_assert(SEL_W == 8, "four row blocks times two column halves");
    constexpr int GPP = SW / 16;                 // K groups per pair
    constexpr int PPM = 4 / GPP;                 // pairs per product
    constexpr int NQD = SW / 4;                  // four-term pieces per row
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const uint32_t p = md.p;
    // P's digits: thread = (column, four terms); C's: (slot, four terms)
    for (int u = tid; u < 2 * PB * NQD; u += SEL_T) {
        const int col = u / NQD, k = u % NQD;
        uint32_t w[ND] = {};
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const uint32_t d4 = digits_n<ND>(sh.X[sh.qs[4 * k + b]][col], p, dg.hi);
#pragma unroll
            for (int i = 0; i < ND; ++i) w[i] |= ((d4 >> (8 * i)) & 0xFFu) << (8 * b);
        }
#pragma unroll
        for (int i = 0; i < ND; ++i) *(uint32_t *)&sh.Pd[i][col][4 * k] = w[i];
    }
    for (int u = tid; u < PB * NQD; u += SEL_T) {
        const int s = u / NQD, k = u % NQD;
        uint32_t wc[ND] = {};
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const uint32_t d4 = digits_n<ND>(sh.C[s][4 * k + b], p, dg.hi);
#pragma unroll
            for (int i = 0; i < ND; ++i) wc[i] |= ((d4 >> (8 * i)) & 0xFFu) << (8 * b);
        }
#pragma unroll
        for (int i = 0; i < ND; ++i) *(uint32_t *)&sh.Cd[i][s][4 * k] = wc[i];
    }
    __syncthreads();
    const int rb = wv & 3, lr = lane & 15, g = lane >> 4;
    // product m of digit sum t: this lane's K group g holds terms
    // 16 (g % GPP) .. of the pair (i, t - i), i = lowest + PPM m + g / GPP
    constexpr int NM = (ND + PPM - 1) / PPM;     // products per digit sum, at most
    const int ho = 16 * (g % GPP);
#pragma unroll
    for (int cbi = 0; cbi < 4; ++cbi) {
        const int cb16 = (wv >> 2) * 4 + cbi;
        v4i acc[2 * ND - 1];
#pragma unroll
        for (int t = 0; t < 2 * ND - 1; ++t) {
            const int ilo = t - (ND - 1) > 0 ? t - (ND - 1) : 0, ihi = t < ND - 1 ? t : ND - 1;
            acc[t] = v4i{0, 0, 0, 0};
#pragma unroll
            for (int m = 0; m < NM; ++m) {
                if (ilo + PPM * m > ihi) continue;  // no pair left for this product
                const int i = ilo + PPM * m + g / GPP;
                const bool ok = i <= ihi;
                const v4i A = ok ? *(const v4i *)&sh.Cd[ok ? i : 0][rb * 16 + lr][ho] : v4i{0, 0, 0, 0};
                const v4i B = ok ? *(const v4i *)&sh.Pd[ok ? t - i : 0][cb16 * 16 + lr][ho] : v4i{0, 0, 0, 0};
                acc[t] = __builtin_amdgcn_mfma_i32_16x16x64_i8(A, B, acc[t], 0, 0, 0);
            }
        }
        // result register e of lane l: slot 16 rb + 4 (l / 16) + e, column l % 16
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int s = rb * 16 + 4 * g + e, col = cb16 * 16 + lr;
            const uint32_t dx = md.mul(sh.D[s], sh.X[s][col]);
            uint32_t r;
            if constexpr (ND == 2) {
                const int32_t ip = (int32_t)p;
                int32_t v = (int32_t)dx + digits_balanced2(acc, e, ip, dg.pinvf);
                v += v < 0 ? ip : 0;
                v -= v >= ip ? ip : 0;
                r = (uint32_t)v;
            } else {
                r = dx + digits_combine<ND>(acc, e, md, dg);
                r = r >= p ? r - p : r;
            }
            sh.X[s][col] = r;
        }
    }
}

// The pivot chain over the sub-panel [cb, cb + SW) of the tile rows um,
// run by every wave (lane = slot): column j of the sub-panel and its
// combination coefficient belong to wave j % SEL_W, the row scale D is kept
// by every wave.  Step c: the owner of column c finds the pivot (the lowest
// unselected tile row nonzero there: the columns left of c are zero in
// every unselected row) from every slot's entry of column c, which its
// owner posts as the step's record; every wave, once each of its lanes
// reads the record with the step's tag, finds the same pivot and
// eliminates its own columns, every other slot x becoming a x - x[c] *
// pivot row.  No barrier per step: the waves run apart, each as far as the
// records posted so far take it, and the owner of step c + 1 posts it
// before its other columns of step c.  Records the
// pivots as E rows kk, kk + 1, ... and leaves the slots' combinations in C
// (column c of the sub-panel: coefficient of the row that pivots there,
// zero if none), D and qs.  Returns the number of pivots (the same in
// every thread); um loses the selected slots.  X is only read: the apply
// that follows writes every column.
constexpr int NCW = SW / SEL_W;  // sub-panel columns per wave
static_assert(SW % SEL_W == 0, "sub-panel columns split evenly over the waves");

// one step of the pivot chain on a wave's own columns wv, wv + SEL_W, ...:
// column c, pivot slot is, pivot value av, this slot's entry f of column c.  Every other slot x becomes av x - f * pivot row (its entries
// right of c) and its combination coefficients follow; the pivot's slot
// keeps its own.
template <int P16>
__device__ __forceinline__ void chain_step(uint32_t (&v)[NCW], uint32_t (&cc)[NCW], uint32_t &dd,
                                           uint32_t &pm, uint64_t &um, int &np, int c, int is,
                                           uint32_t av, uint32_t f, int wv, int lane, const Mod &md,
                                           float pinv) {
    const uint32_t p = md.p;
    const uint32_t dp = (uint32_t)__builtin_amdgcn_readlane((int)dd, is);
    const bool me = lane == is;
    const uint32_t nf = p - f;
#pragma unroll
    for (int i = 0; i < NCW; ++i) {
        const int j = wv + SEL_W * i;  // this wave's column (uniform)
        // new values first, one unconditional write of each at the end (a
        // store per branch would be merged through a pointer and put the
        // arrays in scratch memory)
        uint32_t nv = v[i], nc = cc[i];
        if (j > c) {
            const uint32_t pv = (uint32_t)__builtin_amdgcn_readlane((int)v[i], is);
            const uint32_t t = sel_mad<P16>(av, v[i], nf, pv, md, pinv);
            nv = me ? nv : t;
        } else if (j < c) {
            if ((pm >> j) & 1u) {
                const uint32_t pc = (uint32_t)__builtin_amdgcn_readlane((int)cc[i], is);
                const uint32_t t = sel_mad<P16>(av, cc[i], nf, pc, md, pinv);
                nc = me ? nc : t;
            }
        } else {
            const uint32_t t = sel_mad<P16>(dp, nf, 0u, 0u, md, pinv);
            nc = me ? 0u : t;
        }
        v[i] = nv;
        cc[i] = nc;
    }
    const uint32_t ndd = sel_mad<P16>(av, dd, 0u, 0u, md, pinv);
    dd = me ? dd : ndd;
    pm |= 1u << c;
    um &= ~(1ull << is);
    ++np;
}

template <int P16>
__device__ __forceinline__ int sel_chain(SelShared &sh, int cb, int bw, uint64_t &um, int kk,
                                         int32_t c0, uint32_t &pm, const Mod &md, float pinv,
                                         uint32_t gen) {
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    uint32_t v[NCW], cc[NCW];
    uint32_t dd = 1u;
    bool any = false;
#pragma unroll
    for (int i = 0; i < NCW; ++i) {
        const int j = wv + SEL_W * i;
        v[i] = cb + j < bw ? sh.X[lane][cb + j] : 0u;
        cc[i] = 0u;
        any |= v[i] != 0u;
    }
    pm = 0;  // sub-panel columns with a pivot (uniform)
    // no unselected tile row is nonzero on the sub-panel: no pivot in it
    const uint64_t nzb = um & (uint64_t)__ballot(any);
    if (lane == 0) sh.nzw[wv] = nzb != 0;
    __syncthreads();
    int anyz = 0;
#pragma unroll
    for (int w = 0; w < SEL_W; ++w) anyz |= sh.nzw[w];
    if (!__builtin_amdgcn_readfirstlane(anyz)) return 0;
    // step cn's record: every slot's entry of column cn after the steps
    // before it, each lane's one 64-bit word carrying the tag of this chain
    // and step (never zero; the records start zeroed)
    const uint32_t g64 = gen << 6;
    auto post = [&](int cn, uint32_t val) {
        __hip_atomic_store(&sh.f[cn][lane], ((uint64_t)(g64 + cn + 1) << 32) | val,
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    };
    if (wv == 0) post(0, v[0]);
    int np = 0;
#pragma unroll
    for (int c = 0; c < SW; ++c) {
        const int own = c % SEL_W, ic = c / SEL_W;
        const int nown = (c + 1) % SEL_W, nic = (c + 1) / SEL_W;  // next step's owner
        uint32_t fv;
        if (wv == own) {
            fv = v[ic];  // column c after steps 0 .. c - 1, which this wave took
        } else {
            // step c's record, complete once every lane's word has its tag:
            // each owner posts its step after taking the earlier ones, so
            // the record comes; the bound only guards against a hang
            uint64_t rec = 0;
            for (int spin = 0; spin < (1 << 24); ++spin) {
                rec = __hip_atomic_load(&sh.f[c][lane], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                if (__ballot((uint32_t)(rec >> 32) == g64 + c + 1) == ~0ull) break;
            }
            fv = (uint32_t)rec;
        }
        // the pivot: the lowest unselected slot nonzero on column c (the
        // columns left of c are zero in every unselected slot)
        const uint64_t ball = um & (uint64_t)__ballot(fv != 0u);
        if (!ball) {
            if (c + 1 < SW && wv == nown) post(c + 1, v[nic]);
            continue;
        }
        const int is = __builtin_amdgcn_readfirstlane(__ffsll((unsigned long long)ball) - 1);
        const uint32_t av = (uint32_t)__builtin_amdgcn_readlane((int)fv, is);
        if (c + 1 < SW && wv == nown) {
            // the next step's column first, so that its record goes out
            // before this wave's other columns of step c
            const uint32_t pv = (uint32_t)__builtin_amdgcn_readlane((int)v[nic], is);
            const uint32_t t = sel_mad<P16>(av, v[nic], md.p - fv, pv, md, pinv);
            post(c + 1, lane == is ? v[nic] : t);
        }
        if (wv == own && lane == 0) {
            sh.qs[c] = is;
            sh.ecol[kk + np] = c0 + cb + c;
            sh.eslot[kk + np] = is;
        }
        chain_step<P16>(v, cc, dd, pm, um, np, c, is, av, fv, wv, lane, md, pinv);
    }
#pragma unroll
    for (int i = 0; i < NCW; ++i) {
        const int j = wv + SEL_W * i;
        sh.C[lane][j] = cc[i];
        if (lane == 0 && !((pm >> j) & 1u)) sh.qs[j] = 0;  // no pivot: coefficient zero
    }
    if (wv == 0) sh.D[lane] = dd;
    return np;
}

// P16: 1 for p < 2^16, 0 for larger p, 2 to decide at run time; A's
// entries are T (uint16_t for p <= 65521 in the MFMA form, else uint32_t)
template <int P16, typename T>
__device__ __forceinline__ void select_body(SelShared &sh, const T *a, int64_t ld,
                                            int32_t n, int32_t c0, int32_t *cand,
                                            int32_t *rowpiv, int32_t *colpiv, Panel *pan,
                                            ListState *st, Mod md, const Digits &dg,
                                            int32_t *hdone) {
    if constexpr (P16 == 2) {
        if (md.p < (1u << 16)) select_body<1, T>(sh, a, ld, n, c0, cand, rowpiv, colpiv, pan, st, md, dg, hdone);
        else select_body<0, T>(sh, a, ld, n, c0, cand, rowpiv, colpiv, pan, st, md, dg, hdone);
        return;
    } else {
    const int tid = threadIdx.x, wv = tid >> 6;
    const uint32_t p = md.p;
    const bool small = p < (1u << 17);  // 65 products of residues fit in 64 bits
    const float pinv = 1.0f / (float)p;
    uint32_t r64 = md.reduce(~0ull) + 1u;  // 2^64 mod p
    if (r64 == p) r64 = 0;
    const int32_t done = __builtin_amdgcn_readfirstlane(st->done);
    if (done) {
        // nothing left: this panel and every later one are empty, which
        // the host learns from hdone (mapped host memory, if given)
        if (tid == 0) {
            pan->k = 0;
            if (hdone) *(volatile int32_t *)hdone = 1;
        }
        return;
    }
    const int32_t head = __builtin_amdgcn_readfirstlane(st->head);
    const int32_t count = __builtin_amdgcn_readfirstlane(st->count);
    const int bw = min(PB, n - c0);
    int kk = 0;  // pivots found (same value in every thread)
    int nzero = 0;
    int32_t pos = head;
    uint64_t pmask = 0;  // slots holding E rows (same value in every thread)
    uint64_t epm = 0;    // panel columns of E's pivots (relative to c0)
    for (int i = tid; i < PB * XW; i += SEL_T) (&sh.X[0][0])[i] = 0u;
    for (int i = tid; i < SW * PB; i += SEL_T) (&sh.f[0][0])[i] = 0ull;
    uint32_t gen = 0;  // chains run so far (tags their step records)
    while (pos < count && kk < bw) {
        const int k0 = kk;  // pivots found before this tile
        const uint64_t freem = ~pmask;
        const int nc = min(PB - kk, count - pos);
        // the tile takes the nc lowest free slots
        uint64_t tmask = freem;
        if (nc < PB - kk) {
            tmask = 0;
            uint64_t f = freem;
            for (int j = 0; j < nc; ++j) {
                const uint64_t low = f & (~f + 1ull);
                tmask |= low;
                f ^= low;
            }
        }
        if (tid < PB && ((tmask >> tid) & 1ull)) {
            const int rk = __popcll(freem & ((1ull << tid) - 1ull));
            sh.srow[tid] = cand[pos + rk];
            sh.spos[tid] = pos + rk;
        }
        __syncthreads();
        // 1. load the tile: panel entries (a thread's loads all in flight at
        // once, a row's 64 columns by one wave), unit transform
        {
            constexpr int NL = PB * PB / SEL_T;
            const int col = tid & (PB - 1), s0 = wv;
            uint32_t vals[NL];
#pragma unroll
            for (int i = 0; i < NL; ++i) {
                const int s = s0 + i * SEL_W;
                vals[i] = (((tmask >> s) & 1ull) && col < bw)
                              ? a[(int64_t)sh.srow[s] * ld + c0 + col] : 0u;
            }
#pragma unroll
            for (int i = 0; i < NL; ++i) {
                const int s = s0 + i * SEL_W;
                if ((tmask >> s) & 1ull) {
                    sh.X[s][col] = vals[i];
                    sh.X[s][PB + col] = col == s ? 1u : 0u;
                }
            }
        }
        __syncthreads();
        // 2. reduce the tile by E (normalised): row s -= sum_t X[s][ecol t] E_t.
        // Columns of E's pivots become zero; the other panel columns now,
        // the transform part once a chain of this tile has found a pivot
        // (a tile without pivots never needs it).
        const int ke = kk;
        bool trpend = false;
        const uint64_t bwm = bw == PB ? ~0ull : ((1ull << bw) - 1ull);
        if (ke > 0) {
            for (int i = tid; i < PB * PB; i += SEL_T) {
                const int s = i >> 6, t = i & 63;
                const uint32_t g = (t < ke && ((tmask >> s) & 1ull)) ? sh.X[s][sh.ecol[t] - c0] : 0u;
                sh.CE[s][t] = g ? p - g : 0u;
            }
            if (tid < PB) sh.qse[tid] = tid < ke ? sh.eslot[tid] : 0;
            __syncthreads();
            if (P16 == 1) sel_apply<PB, 0>(sh, sh.CE, sh.qse, true, ke, bwm & ~epm, false, md, r64);
            else if (small) sel_apply<PB, 1>(sh, sh.CE, sh.qse, true, ke, bwm & ~epm, false, md, r64);
            else sel_apply<PB, 2>(sh, sh.CE, sh.qse, true, ke, bwm & ~epm, false, md, r64);
            for (int i = tid; i < PB * PB; i += SEL_T) {
                const int s = i >> 6, t = i & 63;
                if (t < ke && ((tmask >> s) & 1ull)) sh.X[s][sh.ecol[t] - c0] = 0u;
            }
            __syncthreads();
            trpend = true;
        }
        // 3. the pivots of the tile, sub-panel by sub-panel
        uint64_t um = tmask;
        for (int cb = 0; cb < bw && kk < bw; cb += SW) {
            uint32_t pm;
            const int np = sel_chain<P16>(sh, cb, bw, um, kk, c0, pm, md, pinv, gen++);
            __syncthreads();
            if (np > 0) {
                if (trpend) {
                    if (P16 == 1) sel_apply<PB, 0>(sh, sh.CE, sh.qse, true, ke, 0, true, md, r64);
                    else if (small) sel_apply<PB, 1>(sh, sh.CE, sh.qse, true, ke, 0, true, md, r64);
                    else sel_apply<PB, 2>(sh, sh.CE, sh.qse, true, ke, 0, true, md, r64);
                    __syncthreads();
                    trpend = false;
                }
                kk += np;
                epm |= (uint64_t)pm << cb;
                // on the matrix cores in the MFMA form (digits given), on
                // the integer ALU in the plain form
                switch (dg.nd) {
                    case 1: sel_apply_mfma<1>(sh, md, dg); break;
                    case 2: sel_apply_mfma<2>(sh, md, dg); break;
                    case 3: sel_apply_mfma<3>(sh, md, dg); break;
                    case 4: sel_apply_mfma<4>(sh, md, dg); break;
                    default:
                        if (P16 == 1) sel_apply<SW, 0>(sh, sh.C, sh.qs, false, SW, bwm, true, md, r64);
                        else if (small) sel_apply<SW, 1>(sh, sh.C, sh.qs, false, SW, bwm, true, md, r64);
                        else sel_apply<SW, 2>(sh, sh.C, sh.qs, false, SW, bwm, true, md, r64);
                }
                __syncthreads();
            }
        }
        pmask |= tmask & ~um;
        // 4. normalise E by its pivot values (inverses in parallel)
        if (kk > k0) {
            if (tid < kk) sh.einv[tid] = inv_mod(sh.X[sh.eslot[tid]][sh.ecol[tid] - c0], p);
            __syncthreads();
            for (int i = tid; i < kk * 2 * PB; i += SEL_T) {
                const int t = i >> 7, col = i & (2 * PB - 1);
                const int s = sh.eslot[t];
                sh.X[s][col] = md.mul(sh.X[s][col], sh.einv[t]);
            }
            __syncthreads();
        }
        // the sweep went over every column: the rows of this tile that were
        // not selected are zero on the panel
        if (kk < bw) nzero += __popcll(um);
        pos += nc;
    }
    __syncthreads();
    if (tid < kk) {
        const int s = sh.eslot[tid];
        const int32_t r = sh.srow[s];
        sh.erow[tid] = r;
        sh.epos[tid] = sh.spos[s];
        rowpiv[r] = sh.ecol[tid];
        colpiv[sh.ecol[tid]] = r;
        pan->erow[tid] = r;
        pan->ecol[tid] = sh.ecol[tid];
    }
    for (int i = tid; i < PB * PB; i += SEL_T) {
        const int t = i >> 6, sl = i & 63;
        pan->T[i] = (t < kk && sl < kk) ? sh.X[sh.eslot[t]][PB + sh.eslot[sl]] : 0u;
    }
    // move the selected rows to cand[head, head + kk): the j-th selected row
    // further back takes the j-th unselected row of the front, all at once
    if (tid < PB) sh.fr[tid] = 0;
    __syncthreads();
    const int lane = tid & 63;
    const int dpos = tid < kk ? sh.epos[tid] - head : PB;
    if (dpos < kk) sh.fr[dpos] = 1;
    __syncthreads();
    if (wv == 0) {
        const uint64_t lt = (1ull << lane) - 1ull;
        const bool freef = lane < kk && !sh.fr[lane];
        const uint64_t bf = (uint64_t)__ballot(freef);
        if (freef) sh.fpos[__popcll(bf & lt)] = lane;
        const bool back = lane < kk && dpos >= kk;
        const uint64_t bb = (uint64_t)__ballot(back);
        sh.mvj[lane] = __popcll(bb & lt);
    }
    __syncthreads();
    const bool back = tid < kk && dpos >= kk;
    if (back) sh.mvv[tid] = cand[head + sh.fpos[sh.mvj[tid]]];
    __syncthreads();  // every read of the front is done before it is written
    if (back) cand[head + dpos] = sh.mvv[tid];
    if (tid < kk) cand[head + tid] = sh.erow[tid];
    if (tid == 0) {
        pan->k = kk;
        st->head = head + kk;
        st->nzero += nzero;
        if (head + kk >= count) st->done = 1;
    }
    }
}

template <int P16, typename T>
__global__ __launch_bounds__(SEL_T) void k_panel_select(const T *a, int64_t ld, int32_t n,
                                                        int32_t c0, int32_t *cand,
                                                        int32_t *rowpiv, int32_t *colpiv,
                                                        Panel *pan, ListState *st, Mod md,
                                                        Digits dg, int32_t *hdone,
                                                        LoopSync ls = LoopSync(), int32_t j = 0) {
    __shared__ SelShared sh;
    // with ls.w: panel j's select has begun
    if (ls.w && threadIdx.x == 0) ls_post(ls, LS_SEL, ls.s0 + j);
    select_body<P16, T>(sh, a, ld, n, c0, cand, rowpiv, colpiv, pan, st, md, dg, hdone);
}

// ---------------------------------------------------------------- prune
//
// After a panel's update, a candidate row whose columns [c1, n) are all
// zero is zero for good (its columns left of c1 are zero already): when the
// selects since the last prune met at least a tile's worth of rows that
// were zero on their panel, such rows are dropped from the candidate list.
template <typename T>
__device__ __forceinline__ void prune_flags_body(const T *a, int64_t ld, int32_t n,
                                                 int32_t c1, const int32_t *cand,
                                                 const ListState *st, int32_t *alive,
                                                 int32_t first, int32_t step) {
    const int32_t nzero = __builtin_amdgcn_readfirstlane(st->nzero);
    const int32_t done = __builtin_amdgcn_readfirstlane(st->done);
    if (nzero < PB || done) return;
    const int32_t head = __builtin_amdgcn_readfirstlane(st->head);
    const int32_t count = __builtin_amdgcn_readfirstlane(st->count);
    const int lane = threadIdx.x & 63;
    for (int32_t w = first; head + w < count; w += step) {
        const T *row = a + (int64_t)cand[head + w] * ld;
        int nz = 0;  // uniform over the wave
        for (int32_t cb = c1; cb < n; cb += 64) {
            const int32_t c = cb + lane;
            if (__any(c < n && row[c] != 0u)) {
                nz = 1;
                break;
            }
        }
        alive[w] = nz;
    }
}

__device__ __forceinline__ void prune_compact_body(int32_t *cand, ListState *st,
                                                   const int32_t *alive, bool coherent = false) {
    __shared__ int32_t wsum[16];
    const int32_t nzero = __builtin_amdgcn_readfirstlane(st->nzero);
    const int32_t done = __builtin_amdgcn_readfirstlane(st->done);
    if (nzero < PB || done) return;
    const int32_t head = __builtin_amdgcn_readfirstlane(st->head);
    const int32_t count = __builtin_amdgcn_readfirstlane(st->count);
    const int tid = threadIdx.x, wv = tid >> 6, nw = (int)(blockDim.x >> 6);
    int32_t out = head;
    // stable in-place compaction: writes never pass the reads of later chunks
    for (int32_t base = head; base < count; base += (int32_t)blockDim.x) {
        const int32_t i = base + tid;
        const int f = i < count ? (coherent ? __hip_atomic_load(&alive[i - head], __ATOMIC_RELAXED,
                                                                  __HIP_MEMORY_SCOPE_AGENT)
                                            : alive[i - head])
                                : 0;
        const int32_t v = i < count ? cand[i] : 0;
        const int incl = wave_incl_scan(f);
        wsum[wv] = __shfl(incl, 63, 64);  // every lane stores the wave total
        __syncthreads();
        int off = 0, tot = 0;
        for (int q = 0; q < nw; ++q) {
            const int ws = wsum[q];
            if (q < wv) off += ws;
            tot += ws;
        }
        if (f) cand[out + off + incl - 1] = v;
        out += tot;
        __syncthreads();
    }
    if (tid == 0) {
        st->count = out;
        st->nzero = 0;
        if (out <= head) st->done = 1;
    }
}

template <typename T>
__global__ void k_prune_flags(const T *a, int64_t ld, int32_t n, int32_t c1,
                              const int32_t *cand, const ListState *st, int32_t *alive) {
    prune_flags_body(a, ld, n, c1, cand, st, alive,
                     blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6),
                     gridDim.x * (blockDim.x >> 6));
}

__global__ __launch_bounds__(256) void k_prune_compact(int32_t *cand, ListState *st,
                                                       const int32_t *alive) {
    prune_compact_body(cand, st, alive);
}

// both steps in one launch: every block writes its rows' flags, and the last
// block to finish compacts the list (the flags read through to L2)
// block b of nb prune blocks (256 threads each)
template <typename T>
__device__ __forceinline__ void prune_body(const T *a, int64_t ld, int32_t n, int32_t c1,
                                           int32_t *cand, ListState *st, int32_t *alive,
                                           int32_t b, int32_t nb) {
    __shared__ int32_t last;
    const int32_t nzero = __builtin_amdgcn_readfirstlane(st->nzero);
    const int32_t done = __builtin_amdgcn_readfirstlane(st->done);
    if (nzero < PB || done) return;  // the same in every block: st changes below, after all checked
    prune_flags_body(a, ld, n, c1, cand, st, alive, b * (blockDim.x >> 6) + (threadIdx.x >> 6),
                     nb * (blockDim.x >> 6));
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) last = atomicAdd(&st->pcount, 1) == nb - 1;
    __syncthreads();
    if (!last) return;
    __threadfence();
    prune_compact_body(cand, st, alive, true);
    if (threadIdx.x == 0) st->pcount = 0;
}

template <typename T>
__global__ __launch_bounds__(256) void k_prune(const T *a, int64_t ld, int32_t n, int32_t c1,
                                               int32_t *cand, ListState *st, int32_t *alive) {
    prune_body(a, ld, n, c1, cand, st, alive, (int32_t)blockIdx.x, (int32_t)gridDim.x);
}

// ---------------------------------------------------------------- rp

// Rp[t][j] = sum_s T[t][s] A[erow[s]][j] for j >= c0: a block covers 64
// columns, each wave 16 of the rows t; the 64 x 64 panel of the selected
// rows is staged in LDS.  Products are summed in 128 bits (lo + carries)
// and reduced once per output.
__device__ __forceinline__ void rp_body(uint32_t *a, int64_t ld, int32_t n, int32_t c0,
                                        const Panel *pan, uint32_t *rp, int64_t ldr, Mod md,
                                        uint32_t r64, int32_t tile, const Digits &dg) {
    const int kk = __builtin_amdgcn_readfirstlane(pan->k);
    if (kk == 0) return;
    __shared__ uint32_t T[PB][PB + 1];
    __shared__ uint32_t V[PB][PB + 1];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int32_t jb = c0 + tile * 64;
    for (int i = tid; i < PB * PB; i += 256) T[i >> 6][i & 63] = pan->T[i];
    for (int i = tid; i < PB * PB; i += 256) {
        const int s_ = i >> 6, c = i & 63;
        V[s_][c] = (s_ < kk && jb + c < n) ? a[(int64_t)pan->erow[s_] * ld + jb + c] : 0u;
    }
    __syncthreads();
    const int32_t j = jb + lane;
    uint32_t res[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int t = 16 * wv + q;
        uint64_t lo = 0;
        uint32_t hi = 0;
        for (int s_ = 0; s_ < kk; ++s_) {
            const uint64_t pr = (uint64_t)T[t][s_] * (uint64_t)V[s_][lane];
            const uint64_t o = lo;
            lo += pr;
            hi += lo < o;
        }
        uint32_t r = md.reduce(lo);
        if (hi) r = (uint32_t)((r + (uint64_t)md.mul(hi, r64)) % md.p);
        res[q] = r;
    }
    __syncthreads();  // all reads of the selected rows are done before any write
    if (j < n) {
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int t = 16 * wv + q;
            if (t < kk) {
                rp[(int64_t)t * ldr + (j - c0)] = res[q];
                a[(int64_t)pan->erow[t] * ld + j] = res[q];
            }
        }
    }
    // digits of Rp^T: column j - c0, rows t = 16 wv .. 16 wv + 15 (zero
    // beyond k), one 16-byte store per digit
    if (dg.nd > 0 && j < n) {
        for (int i = 0; i < dg.nd; ++i) {
            uint32_t wd[4];
#pragma unroll
            for (int q4 = 0; q4 < 4; ++q4) {
                uint32_t v = 0;
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    const int q = 4 * q4 + b;
                    const uint32_t r = (16 * wv + q) < kk ? res[q] : 0u;
                    v |= digit_word(r, md.p, dg.hi, i) << (8 * b);
                }
                wd[q4] = v;
            }
            uint4 *dst = (uint4 *)(dg.R + i * dg.rplane + (int64_t)(j - c0) * 64 + 16 * wv);
            *dst = make_uint4(wd[0], wd[1], wd[2], wd[3]);
        }
    }
}

// The same product over the columns [jlo, jhi) (jlo >= c0) in narrower
// blocks: a block covers RC columns and every row t (a thread: one column,
// RC / 4 rows), so the blocks that read a column of the selected rows are
// the ones that write it, and a launch over a few columns still spreads
// over many blocks.  KIND 0 (p < 2^16): products below 2^32 from 24-bit
// multiplies, summed in 64 bits; otherwise 64-bit products with a carry.
template <int RC, int KIND, typename T>
__global__ __launch_bounds__(256) void k_panel_rp(T *a, int64_t ld, int32_t c0, int32_t jlo,
                                                  int32_t jhi, const Panel *pan, uint32_t *rp,
                                                  int64_t ldr, Mod md, uint32_t r64, Digits dg) {
    constexpr int NR = RC / 4;  // rows per thread
    const int kk = __builtin_amdgcn_readfirstlane(pan->k);
    if (kk == 0) return;
    __shared__ uint32_t Ts[PB][PB + 1];
    __shared__ uint32_t V[PB][RC + 1];
    const int tid = threadIdx.x;
    const int32_t jb = jlo + (int32_t)blockIdx.x * RC;
    for (int i = tid; i < PB * PB; i += 256) Ts[i >> 6][i & 63] = pan->T[i];
    for (int i = tid; i < PB * RC; i += 256) {
        const int s_ = i / RC, c = i % RC;
        V[s_][c] = (s_ < kk && jb + c < jhi) ? a[(int64_t)pan->erow[s_] * ld + jb + c] : 0u;
    }
    __syncthreads();
    const int jj = tid % RC, tg = tid / RC;  // column, rows NR tg .. NR tg + NR - 1
    const int32_t j = jb + jj;
    uint32_t res[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) {
        const int t = NR * tg + q;
        uint64_t lo = 0;
        uint32_t hi = 0;
        for (int s_ = 0; s_ < kk; ++s_) {
            if (KIND == 0) {
                // __umul24 returns an int: the low 32 bits go through uint32_t
                lo += (uint64_t)(uint32_t)__umul24(Ts[t][s_], V[s_][jj]);
            } else {
                const uint64_t pr = (uint64_t)Ts[t][s_] * (uint64_t)V[s_][jj];
                const uint64_t o = lo;
                lo += pr;
                hi += lo < o;
            }
        }
        uint32_t r = md.reduce(lo);
        if (KIND != 0 && hi) r = (uint32_t)((r + (uint64_t)md.mul(hi, r64)) % md.p);
        res[q] = r;
    }
    __syncthreads();  // all reads of the selected rows are done before any write
    if (j >= jhi) return;
#pragma unroll
    for (int q = 0; q < NR; ++q) {
        const int t = NR * tg + q;
        if (t < kk) {
            rp[(int64_t)t * ldr + (j - c0)] = res[q];
            a[(int64_t)pan->erow[t] * ld + j] = (T)res[q];
        }
    }
    // digits of Rp^T: column j - c0, rows NR tg .. NR tg + NR - 1 (zero
    // beyond k): one 4-byte store per digit for 4 rows, else byte stores
    for (int i = 0; i < dg.nd; ++i) {
        int8_t *dst = dg.R + i * dg.rplane + (int64_t)(j - c0) * 64 + NR * tg;
        if constexpr (NR % 4 == 0) {
#pragma unroll
            for (int b4 = 0; b4 < NR; b4 += 4) {
                uint32_t v = 0;
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    const uint32_t r = (NR * tg + b4 + b) < kk ? res[b4 + b] : 0u;
                    v |= digit_word(r, md.p, dg.hi, i) << (8 * b);
                }
                *(uint32_t *)(dst + b4) = v;
            }
        } else {
#pragma unroll
            for (int b = 0; b < NR; ++b) {
                const uint32_t r = (NR * tg + b) < kk ? res[b] : 0u;
                dst[b] = (int8_t)(uint8_t)digit_word(r, md.p, dg.hi, i);
            }
        }
    }
}

// Rp over [jlo, jhi): RC = 4 (a block per 4 columns) for the few columns
// the next select waits for, RC = 16 for the rest
template <typename T>
void launch_rp(hipStream_t s, bool narrow, T *a, int64_t ld, int32_t c0, int32_t jlo,
               int32_t jhi, const Panel *pan, uint32_t *rp, int64_t ldr, const Mod &md,
               uint32_t r64, const Digits &dg) {
    if (jhi <= jlo) return;
    const bool k0 = md.p < (1u << 16);
    const int rc = narrow ? 4 : 16;
    int64_t g = (jhi - jlo + rc - 1) / rc;
    if (narrow) {
        if (k0) k_panel_rp<4, 0, T><<<(unsigned)g, 256, 0, s>>>(a, ld, c0, jlo, jhi, pan, rp, ldr, md, r64, dg);
        else k_panel_rp<4, 1, T><<<(unsigned)g, 256, 0, s>>>(a, ld, c0, jlo, jhi, pan, rp, ldr, md, r64, dg);
    } else {
        if (k0) k_panel_rp<16, 0, T><<<(unsigned)g, 256, 0, s>>>(a, ld, c0, jlo, jhi, pan, rp, ldr, md, r64, dg);
        else k_panel_rp<16, 1, T><<<(unsigned)g, 256, 0, s>>>(a, ld, c0, jlo, jhi, pan, rp, ldr, md, r64, dg);
    }
}

// ------------------------------------------------------------ gather

// L[r][t] = A[r][ecol[t]] for rows that are not this panel's pivots;
// active[r] = 1 if any of them is nonzero
template <typename T>
__device__ __forceinline__ void gather_body(const T *a, int64_t ld, int32_t m, int32_t c0,
                                            int32_t bw, const int32_t *rowpiv, const Panel *pan,
                                            uint32_t *L, int32_t *active, int32_t first,
                                            int32_t step, uint32_t p, const Digits &dg) {
    const int kk = __builtin_amdgcn_readfirstlane(pan->k);
    if (kk == 0) return;
    const int lane = threadIdx.x & 63;
    for (int32_t r = first; r < m; r += step) {
        const int32_t rp = rowpiv[r];
        const bool cur = rp >= c0 && rp < c0 + bw;
        uint32_t v = 0;
        if (!cur && lane < kk) v = a[(int64_t)r * ld + pan->ecol[lane]];
        L[(int64_t)r * PB + lane] = v;
        for (int i = 0; i < dg.nd; ++i)
            dg.L[i * dg.lplane + (int64_t)r * PB + lane] = (int8_t)digit_word(v, p, dg.hi, i);
        active[r] = __any(v != 0) ? 1 : 0;  // every lane stores the same flag
    }
}

template <typename T>
__global__ void k_panel_gather(const T *a, int64_t ld, int32_t m, int32_t c0, int32_t bw,
                               const int32_t *rowpiv, const Panel *pan, uint32_t *L,
                               int32_t *active, uint32_t p, Digits dg) {
    gather_body(a, ld, m, c0, bw, rowpiv, pan, L, active,
                blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6), gridDim.x * (blockDim.x >> 6),
                p, dg);
}

// ------------------------------------------------- update (integer ALU)

// block: 256 columns x RT rows; A[r][j] -= sum_t L[r][t] Rp[t][j] for the
// columns j in [cb, ce)
constexpr int RT = 16;

template <bool BIG>
__device__ __forceinline__ void update_plain_body(uint32_t *a, int64_t ld, int32_t m, int32_t cb,
                                                  int32_t ce, int32_t c0, const Panel *pan,
                                                  const uint32_t *rp, int64_t ldr,
                                                  const uint32_t *L, const int32_t *active,
                                                  Mod md, uint32_t r64, int32_t bx, int32_t by) {
    const int kk = __builtin_amdgcn_readfirstlane(pan->k);
    if (kk == 0) return;
    __shared__ uint32_t Ls[RT][PB];
    __shared__ int32_t act[RT];
    const int32_t r0 = by * RT;
    for (int i = threadIdx.x; i < RT * PB; i += 256) {
        const int rr = i / PB, t = i % PB;
        Ls[rr][t] = (r0 + rr < m) ? L[(int64_t)(r0 + rr) * PB + t] : 0u;
    }
    if ((int)threadIdx.x < RT) act[threadIdx.x] = (r0 + (int32_t)threadIdx.x < m) ? active[r0 + threadIdx.x] : 0;
    __syncthreads();
    int anyact = 0;
#pragma unroll
    for (int rr = 0; rr < RT; ++rr) anyact |= act[rr];
    if (!anyact) return;
    const int32_t j = cb + bx * 256 + threadIdx.x;
    if (j >= ce) return;
    uint32_t rv[PB];
#pragma unroll
    for (int t = 0; t < PB; ++t) rv[t] = t < kk ? rp[(int64_t)t * ldr + (j - c0)] : 0u;
    for (int rr = 0; rr < RT; ++rr) {
        if (!act[rr]) continue;
        uint64_t lo = 0;
        uint32_t hi = 0;
#pragma unroll
        for (int t = 0; t < PB; ++t) {
            const uint64_t pr = (uint64_t)Ls[rr][t] * (uint64_t)rv[t];
            if (BIG) {
                const uint64_t o = lo;
                lo += pr;
                hi += lo < o;
            } else {
                lo += pr;
            }
        }
        uint32_t s = md.reduce(lo);
        if (BIG && hi) s = (uint32_t)((s + (uint64_t)md.mul(hi, r64)) % md.p);
        uint32_t *x = a + (int64_t)(r0 + rr) * ld + j;
        *x = sub_mod(*x, s, md.p);
    }
}

template <bool BIG>
__global__ __launch_bounds__(256) void k_update_plain(uint32_t *a, int64_t ld, int32_t m,
                                                      int32_t cb, int32_t ce, int32_t c0,
                                                      const Panel *pan,
                                                      const uint32_t *rp, int64_t ldr,
                                                      const uint32_t *L, const int32_t *active,
                                                      Mod md, uint32_t r64) {
    update_plain_body<BIG>(a, ld, m, cb, ce, c0, pan, rp, ldr, L, active, md, r64, blockIdx.x,
                           blockIdx.y);
}

// ------------------------------------------------- update (MFMA int8)

// Block: 4 independent waves, 64 rows x 64 columns; each wave a 32 x 32
// Products v_mfma_i32_16x16x64_i8 (A 16 x 64, B 64 x 16, C 16 x 16; K =
// 64 pivot slots).  With L = sum_i 256^i Li and Rp = sum_j 256^j Rj in
// signed digits, L Rp = sum_s 256^s Cs with Cs = sum_{i+j=s} Li Rj: nd^2
// products into 2 nd - 1 int32 accumulators, each below
// min(s + 1, 2 nd - 1 - s) 256 128^2 <= 2^24 over up to 256 slots and so
// exact.  Operand layout per lane l: A row l % 16, bytes 16 (l / 16) .. + 15
// of the 64 slots; B column l % 16, the same slots; result register e of
// lane l is C[4 (l / 16) + e][l % 16] (measured on gfx950 with an in-bounds
// probe of the builtin).
// A block tile of 64 x 64: four waves of 32 x 32 (RS x CS products of
// 16 x 16), over KC chunks of 64 slots from chunk dg.k0.  The operands are
// 16-byte loads of the digit planes (L2-resident); used for K = 64.
template <int ND, int KC, int RS = 2, int CS = 2, typename T = uint32_t>
__device__ __forceinline__ void update_i8_body(T *a, int64_t ld, int32_t m, int32_t cb,
                                               int32_t ce, int32_t c0, const Panel *pan,
                                               const int32_t *active, Mod md, const Digits &dg,
                                               int32_t bx, int32_t by, int32_t gb = INT_MAX,
                                               int32_t gsz = 0) {
    const int kk = __builtin_amdgcn_readfirstlane(pan->k);
    if (kk == 0) return;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    // the block's tile is 32 RS x 32 CS, each of its four waves 16 RS x 16 CS;
    // block tiles from column gb on skip the gsz columns of a gap
    const int32_t r0 = by * (32 * RS) + (wv >> 1) * (16 * RS);
    int32_t qb = cb + bx * (32 * CS);
    if (qb >= gb) qb += gsz;
    const int32_t q0 = qb + (wv & 1) * (16 * CS);
    // rows of this wave with a nonzero multiplier (wave-uniform exit)
    const bool act = lane < 16 * RS && r0 + lane < m && active[r0 + lane] != 0;
    if (!__any(act) || q0 >= ce) return;
    const int lr = lane & 15, lk = lane >> 4;
    // this lane's values of A (rows r0 + 16 rs + 4 (l / 16) + e, columns
    // q0 + 16 cs + l % 16): for nd <= 2 loaded before the products so that
    // their latency hides behind them; for nd >= 3 after them, where the
    // registers would cost a wave per SIMD
    constexpr bool PRE = ND <= 2;
    uint32_t xv[RS][4][CS];
    bool xa[RS][4];
#pragma unroll
    for (int rs = 0; rs < RS; ++rs)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int32_t row = r0 + 16 * rs + 4 * lk + e;
            xa[rs][e] = row < m && active[row] != 0;
            if constexpr (PRE) {
#pragma unroll
                for (int cs = 0; cs < CS; ++cs) {
                    const int32_t col = q0 + 16 * cs + lr;
                    xv[rs][e][cs] = (xa[rs][e] && col < ce) ? a[(int64_t)row * ld + col] : 0u;
                }
            }
        }
    v4i acc[RS][CS][2 * ND - 1];
#pragma unroll
    for (int rs = 0; rs < RS; ++rs)
#pragma unroll
        for (int cs = 0; cs < CS; ++cs)
#pragma unroll
            for (int q = 0; q < 2 * ND - 1; ++q) acc[rs][cs][q] = v4i{0, 0, 0, 0};
#pragma unroll
    for (int kc = 0; kc < KC; ++kc) {
        const int8_t *lb = dg.L + (dg.k0 + kc) * dg.lchunk + 16 * lk;
        const int8_t *rb = dg.R + (dg.k0 + kc) * dg.rchunk + 16 * lk;
        v4i A[ND][RS], B[ND][CS];
#pragma unroll
        for (int i = 0; i < ND; ++i) {
#pragma unroll
            for (int h = 0; h < RS; ++h) {
                const int32_t row = r0 + 16 * h + lr;
                A[i][h] = row < m ? *(const v4i *)(lb + i * dg.lplane + (int64_t)row * 64)
                                  : v4i{0, 0, 0, 0};
            }
#pragma unroll
            for (int h = 0; h < CS; ++h) {
                const int32_t col = q0 + 16 * h + lr;
                B[i][h] = col < ce ? *(const v4i *)(rb + i * dg.rplane + (int64_t)(col - c0) * 64)
                                   : v4i{0, 0, 0, 0};
            }
        }
#pragma unroll
        for (int i = 0; i < ND; ++i)
#pragma unroll
            for (int j = 0; j < ND; ++j)
#pragma unroll
                for (int rs = 0; rs < RS; ++rs)
#pragma unroll
                    for (int cs = 0; cs < CS; ++cs)
                        acc[rs][cs][i + j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(
                            A[i][rs], B[j][cs], acc[rs][cs][i + j], 0, 0, 0);
    }
#pragma unroll
    for (int rs = 0; rs < RS; ++rs)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            if (!xa[rs][e]) continue;
            const int32_t row = r0 + 16 * rs + 4 * lk + e;
#pragma unroll
            for (int cs = 0; cs < CS; ++cs) {
                const int32_t col = q0 + 16 * cs + lr;
                if (col >= ce) continue;
                T *x = a + (int64_t)row * ld + col;
                const uint32_t xo = PRE ? xv[rs][e][cs] : (uint32_t)*x;
                *x = (T)digits_sub<ND>(xo, acc[rs][cs], e, md, dg);
            }
        }
}

template <int ND, int KC = 1, int RS = 2, int CS = 2, typename T = uint32_t>
__global__ __launch_bounds__(256) void k_update_mfma(T *a, int64_t ld, int32_t m,
                                                     int32_t cb, int32_t ce, int32_t c0,
                                                     const Panel *pan, const int32_t *active,
                                                     Mod md, Digits dg, int32_t gb = INT_MAX,
                                                     int32_t gsz = 0) {
    update_i8_body<ND, KC, RS, CS, T>(a, ld, m, cb, ce, c0, pan, active, md, dg, blockIdx.x,
                                   blockIdx.y, gb, gsz);
}

// K = 256 (four chunks from chunk 0) over wide column ranges.  The block's
// four waves take four strips of 16 RS rows and walk WT_NT column tiles of
// 32 together; each wave keeps its rows' L digits in registers for the
// whole walk.  A tile's Rp digits (nd x 4 chunks x 32 columns x 64 bytes)
// are staged in LDS once for the block, one tile ahead (two buffers, one
// barrier per tile), and the A values of the next tile are loaded while
// the current one is multiplied.
constexpr int WT_NT = 16;   // column tiles per block
constexpr int WT_CS = 2;    // 16-column products per tile

template <int ND, int RS, typename T, int KC, int NT = WT_NT>
__device__ __forceinline__ void update_wide_body(T *a, int64_t ld, int32_t m, int32_t cb,
                                                 int32_t ce, int32_t c0, const Panel *pan,
                                                 const int32_t *active, const Mod &md,
                                                 const Digits &dg, int32_t gb, int32_t gsz,
                                                 int32_t bx, int32_t by) {
    constexpr int CS = WT_CS, TW = 16 * CS;
    constexpr int TB = ND * KC * TW * 64;  // bytes of Rp digits per tile
    constexpr int NP = TB / 16;            // its 16-byte pieces
    constexpr int PPT = (NP + 255) / 256;  // per thread (the last round partial)
    __shared__ __align__(16) int8_t rl[2][TB];
    const int kk = __builtin_amdgcn_readfirstlane(pan->k);
    // tiles walk the columns [cb, ce) without the gap [gb, gb + gsz) (gb - cb
    // a multiple of TW): virtual column v is column v, or v + gsz from gb on
    const int32_t qb = cb + bx * (NT * TW);
    const int32_t vce = ce - gsz;
    auto real = [&](int32_t v) { return v >= gb ? v + gsz : v; };
    if (kk == 0 || qb >= vce) return;  // uniform over the block
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int32_t r0 = (by * 4 + wv) * (16 * RS);
    const bool wact = __any(lane < 16 * RS && r0 + lane < m && active[r0 + lane] != 0);
    const int lr = lane & 15, lk = lane >> 4;
    // piece u 256 + tid of a tile: digit i, chunk kc, column c / 4, bytes
    // 16 (c % 4) .. + 15; the LDS image is [i][kc][column][64]
    auto rload = [&](int32_t q0, v4i (&st)[PPT]) {
#pragma unroll
        for (int u = 0; u < PPT; ++u) {
            const int idx = u * 256 + tid;
            if (NP % 256 == 0 || idx < NP) {
                const int ik = idx / (TW * 4), c = idx % (TW * 4);
                const int i = ik / KC, kc = ik % KC;
                int32_t col = q0 + (c >> 2);
                col = col < ce ? col : ce - 1;
                st[u] = *(const v4i *)(dg.R + i * dg.rplane + kc * dg.rchunk +
                                       (int64_t)(col - c0) * 64 + 16 * (c & 3));
            }
        }
    };
    auto rstore = [&](int buf, const v4i (&st)[PPT]) {
#pragma unroll
        for (int u = 0; u < PPT; ++u)
            if (NP % 256 == 0 || u * 256 + tid < NP) *(v4i *)(rl[buf] + (u * 256 + tid) * 16) = st[u];
    };
    v4i Lr[KC][ND][RS];
    if (wact) {
#pragma unroll
        for (int h = 0; h < RS; ++h) {
            const int32_t row = r0 + 16 * h + lr;
            const int8_t *lp = dg.L + (int64_t)(row < m ? row : m - 1) * 64 + 16 * lk;
#pragma unroll
            for (int kc = 0; kc < KC; ++kc)
#pragma unroll
                for (int i = 0; i < ND; ++i)
                    Lr[kc][i][h] = *(const v4i *)(lp + i * dg.lplane + kc * dg.lchunk);
        }
    }
    bool xa[RS][4];
    T *xp[RS][4];
#pragma unroll
    for (int rs = 0; rs < RS; ++rs)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int32_t row = r0 + 16 * rs + 4 * lk + e;
            xa[rs][e] = row < m && active[row] != 0;
            xp[rs][e] = a + (int64_t)(xa[rs][e] ? row : 0) * ld + lr;
        }
    // A values: columns past ce read the row's last column instead
    uint32_t xv[RS][4][CS], xn[RS][4][CS];
    const int32_t clast = ce - 1 - lr;
    auto loadx = [&](int32_t q0, uint32_t (&x)[RS][4][CS]) {
#pragma unroll
        for (int rs = 0; rs < RS; ++rs)
#pragma unroll
            for (int e = 0; e < 4; ++e)
                if (xa[rs][e]) {
#pragma unroll
                    for (int cs = 0; cs < CS; ++cs) {
                        const int32_t col = q0 + 16 * cs;
                        x[rs][e][cs] = (uint32_t)xp[rs][e][col < clast ? col : clast];
                    }
                }
    };
    int nt = (vce - qb + TW - 1) / TW;
    nt = nt < NT ? nt : NT;
    {
        v4i st[PPT];
        rload(real(qb), st);
        rstore(0, st);
    }
    if (wact) loadx(real(qb), xv);
    __syncthreads();
#pragma unroll 1
    for (int t = 0; t < nt; ++t) {
        const int32_t q0 = real(qb + t * TW), qn = real(qb + (t + 1) * TW);
        const bool more = t + 1 < nt;
        v4i st[PPT];
        if (more) {
            rload(qn, st);
            if (wact) loadx(qn, xn);
        }
        if (wact) {
            const int8_t *lb = rl[t & 1] + lr * 64 + 16 * lk;
            v4i acc[RS][CS][2 * ND - 1];
#pragma unroll
            for (int rs = 0; rs < RS; ++rs)
#pragma unroll
                for (int cs = 0; cs < CS; ++cs)
#pragma unroll
                    for (int q = 0; q < 2 * ND - 1; ++q) acc[rs][cs][q] = v4i{0, 0, 0, 0};
#pragma unroll
            for (int kc = 0; kc < KC; ++kc) {
                v4i B[ND][CS];
#pragma unroll
                for (int i = 0; i < ND; ++i)
#pragma unroll
                    for (int h = 0; h < CS; ++h)
                        B[i][h] = *(const v4i *)(lb + ((i * KC + kc) * TW + 16 * h) * 64);
#pragma unroll
                for (int i = 0; i < ND; ++i)
#pragma unroll
                    for (int j = 0; j < ND; ++j)
#pragma unroll
                        for (int rs = 0; rs < RS; ++rs)
#pragma unroll
                            for (int cs = 0; cs < CS; ++cs)
                                acc[rs][cs][i + j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(
                                    Lr[kc][i][rs], B[j][cs], acc[rs][cs][i + j], 0, 0, 0);
            }
            const bool full = q0 + TW <= ce;
#pragma unroll
            for (int rs = 0; rs < RS; ++rs)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    if (!xa[rs][e]) continue;
                    uint32_t v[CS];
#pragma unroll
                    for (int cs = 0; cs < CS; ++cs)
                        v[cs] = digits_sub<ND>(xv[rs][e][cs], acc[rs][cs], e, md, dg);
                    if (full) {
#pragma unroll
                        for (int cs = 0; cs < CS; ++cs) xp[rs][e][q0 + 16 * cs] = (T)v[cs];
                    } else {
#pragma unroll
                        for (int cs = 0; cs < CS; ++cs)
                            if (q0 + 16 * cs + lr < ce) xp[rs][e][q0 + 16 * cs] = (T)v[cs];
                    }
                }
        }
        if (more) {
            rstore((t + 1) & 1, st);
#pragma unroll
            for (int rs = 0; rs < RS; ++rs)
#pragma unroll
                for (int e = 0; e < 4; ++e)
#pragma unroll
                    for (int cs = 0; cs < CS; ++cs) xv[rs][e][cs] = xn[rs][e][cs];
        }
        // the next tile's Rp digits are visible, and every wave is done
        // with the buffer the tile after it overwrites
        __syncthreads();
    }
}

template <int ND, int RS, typename T = uint32_t, int KC = 4, int NT = WT_NT>
__global__ __launch_bounds__(256) void k_update_wide(T *a, int64_t ld, int32_t m,
                                                     int32_t cb, int32_t ce, int32_t c0,
                                                     const Panel *pan, const int32_t *active,
                                                     Mod md, Digits dg, int32_t gb = INT_MAX,
                                                     int32_t gsz = 0) {
    update_wide_body<ND, RS, T, KC, NT>(a, ld, m, cb, ce, c0, pan, active, md, dg, gb, gsz,
                                        (int32_t)blockIdx.x, (int32_t)blockIdx.y);
}

template <typename T>
void launch_update_i8(int nd, int kc, hipStream_t s, T *a, int64_t ld, int32_t m, int32_t cb,
                      int32_t ce, int32_t c0, const Panel *pan, const int32_t *active, Mod md,
                      const Digits &dg) {
    if (cb >= ce || m <= 0) return;
    // 16-bit storage holds residues of p <= 65521: one or two digits
    constexpr bool W16 = sizeof(T) == 2;
    if (kc == 4) {
        // 64 rows per block for nd <= 2 (32 per wave), 32 rows for nd >= 3
        const int rows = nd <= 2 ? 128 : 64;
        const int cols = WT_NT * 16 * WT_CS;
        dim3 grid((unsigned)((ce - cb + cols - 1) / cols), (unsigned)((m + rows - 1) / rows));
        if (nd == 1) k_update_wide<1, 2, T><<<grid, 256, 0, s>>>(a, ld, m, cb, ce, c0, pan, active, md, dg);
        else if (nd == 2 || W16) k_update_wide<2, 2, T><<<grid, 256, 0, s>>>(a, ld, m, cb, ce, c0, pan, active, md, dg);
        else if constexpr (!W16) {
            if (nd == 3) k_update_wide<3, 1><<<grid, 256, 0, s>>>(a, ld, m, cb, ce, c0, pan, active, md, dg);
            else k_update_wide<4, 1><<<grid, 256, 0, s>>>(a, ld, m, cb, ce, c0, pan, active, md, dg);
        }
        return;
    }
    // one and two digits: 32 x 64 block tiles, as in the panel loop
    const int rows = (nd <= 2 || W16) ? 32 : 64;
    dim3 grid((unsigned)((ce - cb + 63) / 64), (unsigned)((m + rows - 1) / rows));
    if (nd == 1) k_update_mfma<1, 1, 1, 2, T><<<grid, 256, 0, s>>>(a, ld, m, cb, ce, c0, pan, active, md, dg);
    else if (nd == 2 || W16) k_update_mfma<2, 1, 1, 2, T><<<grid, 256, 0, s>>>(a, ld, m, cb, ce, c0, pan, active, md, dg);
    else if constexpr (!W16) {
        if (nd == 3) k_update_mfma<3><<<grid, 256, 0, s>>>(a, ld, m, cb, ce, c0, pan, active, md, dg);
        else k_update_mfma<4><<<grid, 256, 0, s>>>(a, ld, m, cb, ce, c0, pan, active, md, dg);
    }
}
template void launch_update_i8<uint32_t>(int, int, hipStream_t, uint32_t *, int64_t, int32_t,
                                         int32_t, int32_t, int32_t, const Panel *, const int32_t *,
                                         Mod, const Digits &);
template void launch_update_i8<uint16_t>(int, int, hipStream_t, uint16_t *, int64_t, int32_t,
                                         int32_t, int32_t, int32_t, const Panel *, const int32_t *,
                                         Mod, const Digits &);

// ------------------------------------------------------------ output

__global__ void k_out_flags(int32_t n, const int32_t *colpiv, int32_t *flag) {
    for (int32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x)
        flag[c] = colpiv[n - 1 - c] >= 0;  // decreasing lead order
}

__global__ void k_out_list(int32_t n, const int32_t *colpiv, const int32_t *flag,
                           const int32_t *rank, int32_t *orow, int32_t *olead) {
    for (int32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x)
        if (flag[c]) {
            orow[rank[c]] = colpiv[n - 1 - c];
            olead[rank[c]] = n - 1 - c;
        }
}

// rowpiv = colpiv = -1, cand = 0, 1, ..., m - 1, panel state reset
__global__ void k_dense_init(int32_t m, int32_t n, int32_t *rowpiv, int32_t *colpiv,
                             int32_t *cand, Panel *pan, ListState *st) {
    const int64_t tot = (int64_t)m > n ? m : n;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < tot;
         i += (int64_t)gridDim.x * blockDim.x) {
        if (i < m) {
            rowpiv[i] = -1;
            cand[i] = (int32_t)i;
        }
        if (i < n) colpiv[i] = -1;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        pan[0].k = 0;
        pan[1].k = 0;
        st->done = m == 0 ? 1 : 0;
        st->head = 0;
        st->count = m;
        st->nzero = 0;
        st->pcount = 0;
    }
}

// the digit planes of T (Td[i][t][s], t, s < kk; zero elsewhere) and of
// the pivot rows' entries of the columns [q0, min(q0 + 64, qe)), transposed
// (Pd[i][j][s]); erow in LDS.  A thread builds four consecutive s of one row
// t and of one column j at a time (a wave reads 64 consecutive columns of a
// pivot row); every load is in flight before the first is used.
template <int ND, typename T, int NT>
__device__ __forceinline__ void tile_digits(int8_t (*Td)[PB][PB], int8_t (*Pd)[PB][PB],
                                            const Panel *pan, const T *a, int64_t ld,
                                            const int32_t *erow, int32_t q0, int32_t qe, int kk,
                                            uint32_t p, uint32_t hi) {
    const int tid = threadIdx.x;
    constexpr int IT = PB * PB / 4 / NT;  // four-element pieces per thread (NT threads)
    uint32_t tv[IT][4], pv[IT][4];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int u = tid + NT * it;
        // T holds zeros past kk in both indices
#pragma unroll
        for (int b = 0; b < 4; ++b) tv[it][b] = pan->T[(u >> 4) * PB + (u & 15) * 4 + b];
        const int j = u & 63, s4 = (u >> 6) * 4;
#pragma unroll
        for (int b = 0; b < 4; ++b)
            pv[it][b] = q0 + j < qe ? (uint32_t)a[(int64_t)erow[s4 + b] * ld + q0 + j] : 0u;
    }
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int u = tid + NT * it;
        uint32_t wt[ND] = {}, wp[ND] = {};
        const int j = u & 63, s4 = (u >> 6) * 4;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const uint32_t dt = digits_n<ND>(tv[it][b], p, hi);
            const uint32_t dp = digits_n<ND>(s4 + b < kk ? pv[it][b] : 0u, p, hi);
#pragma unroll
            for (int i = 0; i < ND; ++i) {
                wt[i] |= ((dt >> (8 * i)) & 0xFFu) << (8 * b);
                wp[i] |= ((dp >> (8 * i)) & 0xFFu) << (8 * b);
            }
        }
#pragma unroll
        for (int i = 0; i < ND; ++i) {
            *(uint32_t *)&Td[i][u >> 4][(u & 15) * 4] = wt[i];
            *(uint32_t *)&Pd[i][j][s4] = wp[i];
        }
    }
}

// ------------------------------------------------- next panel's share
//
// The main stream's part of a panel step in the MFMA form, one launch for
// what the select of the next panel waits for.  Block b takes the rows
// [64 b, 64 b + 64):
//   1. Rp of the next panel's columns [c1, c2): T (the transform of the
//      select) times the pivot rows' entries there, on the matrix cores in
//      signed base-256 digits (every block computes it; block 0 also leaves
//      it in rp, from where the side stream writes the pivot rows once every
//      block has read them);
//   2. the multipliers of its rows: L[r][t] = A[r][ecol t] (zero for this
//      panel's pivot rows), as their digit planes (the plain multipliers
//      only the plain form reads are not written) and the active flags, as
//      gather_body;
//   3. A[r][c1, c2) -= L[r] Rp for its active rows, on the matrix cores.
// The digits go through LDS: T and the pivot rows' entries (operands of 1),
// Rp^T and L (operands of 3).  The blocks from gn on are the previous
// panel's prune of the candidate list over the columns [c0, n) (prune_body;
// the list is not read here, and a row this launch updates has a nonzero
// multiplier in [c0, c1), which it does not write, so the prune's verdict
// does not depend on the order of the two).
constexpr int NAR_T = 512;  // threads of a k_panel_narrow block (eight waves)

template <int ND, typename T>
__device__ __forceinline__ void narrow_body(T *a, int64_t ld, int32_t m, int32_t c0, int32_t c1,
                                            int32_t c2, const int32_t *rowpiv, const Panel *pan,
                                            int32_t *active, uint32_t *rp, int64_t ldr,
                                            const Mod &md, const Digits &dg) {
    const int kk = __builtin_amdgcn_readfirstlane(pan->k);
    if (kk == 0) return;
    __shared__ __align__(16) int8_t Td[ND][PB][PB];  // T: [digit][t][s]
    __shared__ __align__(16) int8_t Pd[ND][PB][PB];  // pivot rows, transposed: [digit][j][s]
    __shared__ __align__(16) int8_t Rd[ND][PB][PB];  // Rp, transposed: [digit][j][t]
    __shared__ __align__(16) int8_t Ld[ND][PB][PB];  // multipliers: [digit][row][t]
    __shared__ int32_t erow_s[PB], ecol_s[PB], act_s[PB];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int lr = lane & 15, lk = lane >> 4;
    const uint32_t p = md.p;
    const int32_t w2 = c2 - c1;  // columns of the next panel (1 .. 64)
    const int32_t r0 = (int32_t)blockIdx.x * PB;
    // wave wv: the rows 8 wv .. 8 wv + 7 of the block in step 2; in steps
    // 1b and 3 the 16 x 16 tiles of row (or t) block tb and column blocks
    // cb0, cb0 + 1
    const int tb = wv & 3, cb0 = 2 * (wv >> 2);
    // loads that need nothing computed here go out first: the pivot
    // columns of this wave's rows and this lane's entries of step 3 (rows
    // 16 tb + 4 (lane / 16) + e, columns c1 + 16 cb + lane % 16)
    int32_t rpv[8];
#pragma unroll
    for (int h = 0; h < 8; ++h) {
        const int32_t r = r0 + 8 * wv + h;
        rpv[h] = r < m ? rowpiv[r] : 0;
    }
    uint32_t xv[4][2];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int32_t r = r0 + 16 * tb + 4 * lk + e;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int32_t j = 16 * (cb0 + q) + lr;
            xv[e][q] = (r < m && j < w2) ? (uint32_t)a[(int64_t)r * ld + c1 + j] : 0u;
        }
    }
    if (tid < PB) {
        erow_s[tid] = tid < kk ? pan->erow[tid] : 0;
        ecol_s[tid] = tid < kk ? pan->ecol[tid] : 0;
    }
    __syncthreads();
    // 1a. digits of T (a thread: row t, four consecutive s) and of the pivot
    // rows (a thread: column j, four consecutive s; a wave reads 64
    // consecutive columns of one row); every load in flight before the
    // first is used
    tile_digits<ND, T, NAR_T>(Td, Pd, pan, a, ld, erow_s, c1, c1 + w2, kk, p, dg.hi);
    // 2a. this wave's rows' multipliers, their loads in flight together
    uint32_t lv[8];
#pragma unroll
    for (int h = 0; h < 8; ++h) {
        const int32_t r = r0 + 8 * wv + h;
        const bool on = r < m && !(rpv[h] >= c0 && rpv[h] < c1);
        lv[h] = (on && lane < kk) ? (uint32_t)a[(int64_t)r * ld + ecol_s[lane]] : 0u;
    }
    __syncthreads();
    // 1b. Rp^T; result register e of lane l: t = 16 tb + 4 (l / 16) + e,
    // j = 16 cb + l % 16
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int cb = cb0 + q;
        v4i acc[2 * ND - 1];
#pragma unroll
        for (int u = 0; u < 2 * ND - 1; ++u) acc[u] = v4i{0, 0, 0, 0};
#pragma unroll
        for (int i = 0; i < ND; ++i)
#pragma unroll
            for (int jd = 0; jd < ND; ++jd)
                acc[i + jd] = __builtin_amdgcn_mfma_i32_16x16x64_i8(
                    *(const v4i *)&Td[i][16 * tb + lr][16 * lk],
                    *(const v4i *)&Pd[jd][16 * cb + lr][16 * lk], acc[i + jd], 0, 0, 0);
        const int j = 16 * cb + lr;
        uint32_t wd[ND] = {};
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int t = 16 * tb + 4 * lk + e;
            uint32_t r;
            if constexpr (ND == 2) {
                const int32_t v = digits_balanced2(acc, e, (int32_t)p, dg.pinvf);
                r = (uint32_t)(v < 0 ? v + (int32_t)p : v);
            } else {
                r = digits_combine<ND>(acc, e, md, dg);
            }
            r = (t < kk && j < w2) ? r : 0u;
            if (blockIdx.x == 0 && t < kk && j < w2) rp[(int64_t)t * ldr + (c1 + j - c0)] = r;
            const uint32_t d4 = digits_n<ND>(r, p, dg.hi);
#pragma unroll
            for (int i = 0; i < ND; ++i) wd[i] |= ((d4 >> (8 * i)) & 0xFFu) << (8 * e);
        }
#pragma unroll
        for (int i = 0; i < ND; ++i) *(uint32_t *)&Rd[i][j][16 * tb + 4 * lk] = wd[i];
    }
    // 2b. the multipliers (digit planes, LDS) and the active flags
#pragma unroll
    for (int h = 0; h < 8; ++h) {
        const int32_t r = r0 + 8 * wv + h;
        const uint32_t v = lv[h];
        const uint32_t d4 = digits_n<ND>(v, p, dg.hi);
        const bool act = __any(v != 0u) && r < m;
        if (r < m) {
#pragma unroll
            for (int i = 0; i < ND; ++i)
                dg.L[i * dg.lplane + (int64_t)r * PB + lane] = (int8_t)((d4 >> (8 * i)) & 0xFFu);
            active[r] = act ? 1 : 0;
        }
#pragma unroll
        for (int i = 0; i < ND; ++i) Ld[i][8 * wv + h][lane] = (int8_t)((d4 >> (8 * i)) & 0xFFu);
        if (lane == 0) act_s[8 * wv + h] = act ? 1 : 0;
    }
    // 3. A[r][c1, c2) -= L[r] Rp for the active rows
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int cb = cb0 + q;
        v4i acc[2 * ND - 1];
#pragma unroll
        for (int u = 0; u < 2 * ND - 1; ++u) acc[u] = v4i{0, 0, 0, 0};
#pragma unroll
        for (int i = 0; i < ND; ++i)
#pragma unroll
            for (int jd = 0; jd < ND; ++jd)
                acc[i + jd] = __builtin_amdgcn_mfma_i32_16x16x64_i8(
                    *(const v4i *)&Ld[i][16 * tb + lr][16 * lk],
                    *(const v4i *)&Rd[jd][16 * cb + lr][16 * lk], acc[i + jd], 0, 0, 0);
        const int32_t j = 16 * cb + lr;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int32_t r = r0 + 16 * tb + 4 * lk + e;
            if (act_s[16 * tb + 4 * lk + e] && j < w2)
                a[(int64_t)r * ld + c1 + j] = (T)digits_sub<ND>(xv[e][q], acc, e, md, dg);
        }
    }
}

template <int ND, typename T>
__global__ __launch_bounds__(NAR_T) void k_panel_narrow(T *a, int64_t ld, int32_t m, int32_t n,
                                                        int32_t c0, int32_t c1, int32_t c2,
                                                        const int32_t *rowpiv, const Panel *pan,
                                                        int32_t *active, uint32_t *rp,
                                                        int64_t ldr, Mod md, Digits dg,
                                                        int32_t gn, int32_t *cand, ListState *st,
                                                        int32_t *alive, LoopSync ls = LoopSync(),
                                                        int32_t jp = 0) {
    // with ls.w: panel jp's launch has begun, so the select before it is
    // complete (the side stream's Rp of panel jp waits for this)
    if (ls.w && blockIdx.x == 0 && threadIdx.x == 0) ls_post(ls, LS_NAR, ls.s0 + jp);
    if ((int32_t)blockIdx.x >= gn) {
        prune_body(a, ld, n, c0, cand, st, alive, (int32_t)blockIdx.x - gn, (int32_t)gridDim.x - gn);
        return;
    }
    narrow_body<ND, T>(a, ld, m, c0, c1, c2, rowpiv, pan, active, rp, ldr, md, dg);
}

// queued on the main stream after panel j's side stream work: one thread
// waits until that work is complete (the next panel's launch reads it)
__global__ void k_main_wait(LoopSync ls, int32_t j) {
    if (threadIdx.x == 0) ls_wait(ls, LS_SIDE, ls.s0 + j);
}

// Rp of the columns [c0, jhi) in the MFMA form, a block per 64 columns
// [c0 + 64 b, c0 + 64 b + 64): T times the pivot rows' entries there, on
// the matrix cores in signed base-256 digits; writes the pivot rows (in
// reduced form) and the digit planes of Rp^T for the update.  The block at
// column wb (the next panel's, whose Rp k_panel_narrow has left in rp)
// writes the pivot rows there from rp instead: every block of that launch
// has read them by now.
template <int ND, typename T>
__device__ __forceinline__ void rp_mfma_body(T *a, int64_t ld, int32_t c0, int32_t jhi, int32_t wb,
                                             const Panel *pan, const uint32_t *rp, int64_t ldr,
                                             const Mod &md, const Digits &dg, int32_t blk) {
    const int kk = __builtin_amdgcn_readfirstlane(pan->k);
    if (kk == 0) return;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int lr = lane & 15, lk = lane >> 4;
    const uint32_t p = md.p;
    const int32_t q0 = c0 + blk * PB;
    if (q0 == wb) {
        for (int i = tid; i < kk * PB; i += 256) {
            const int t = i >> 6, j = i & 63;
            if (q0 + j < jhi) a[(int64_t)pan->erow[t] * ld + q0 + j] = (T)rp[(int64_t)t * ldr + (q0 + j - c0)];
        }
        return;
    }
    __shared__ __align__(16) int8_t Td[ND][PB][PB];  // T: [digit][t][s]
    __shared__ __align__(16) int8_t Pd[ND][PB][PB];  // pivot rows, transposed: [digit][j][s]
    __shared__ int32_t erow_s[PB];
    if (tid < PB) erow_s[tid] = tid < kk ? pan->erow[tid] : 0;
    __syncthreads();
    tile_digits<ND, T, 256>(Td, Pd, pan, a, ld, erow_s, q0, jhi, kk, p, dg.hi);
    __syncthreads();  // every read of the pivot rows is done before they are written
    // wave wv: rows t of 16 wv .., all 64 columns; register e of lane l:
    // t = 16 wv + 4 (l / 16) + e, column 16 cb + l % 16
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) {
        v4i acc[2 * ND - 1];
#pragma unroll
        for (int q = 0; q < 2 * ND - 1; ++q) acc[q] = v4i{0, 0, 0, 0};
#pragma unroll
        for (int i = 0; i < ND; ++i)
#pragma unroll
            for (int jd = 0; jd < ND; ++jd)
                acc[i + jd] = __builtin_amdgcn_mfma_i32_16x16x64_i8(
                    *(const v4i *)&Td[i][16 * wv + lr][16 * lk],
                    *(const v4i *)&Pd[jd][16 * cb + lr][16 * lk], acc[i + jd], 0, 0, 0);
        const int32_t col = q0 + 16 * cb + lr;
        uint32_t wd[ND] = {};
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int t = 16 * wv + 4 * lk + e;
            uint32_t r;
            if constexpr (ND == 2) {
                const int32_t v = digits_balanced2(acc, e, (int32_t)p, dg.pinvf);
                r = (uint32_t)(v < 0 ? v + (int32_t)p : v);
            } else {
                r = digits_combine<ND>(acc, e, md, dg);
            }
            r = (t < kk && col < jhi) ? r : 0u;
            if (t < kk && col < jhi) a[(int64_t)erow_s[t] * ld + col] = (T)r;
            const uint32_t d4 = digits_n<ND>(r, p, dg.hi);
#pragma unroll
            for (int i = 0; i < ND; ++i) wd[i] |= ((d4 >> (8 * i)) & 0xFFu) << (8 * e);
        }
        if (col < jhi) {
#pragma unroll
            for (int i = 0; i < ND; ++i)
                *(uint32_t *)(dg.R + i * dg.rplane + (int64_t)(col - c0) * 64 + 16 * wv + 4 * lk) = wd[i];
        }
    }
}

template <int ND, typename T>
__global__ __launch_bounds__(256) void k_panel_rp_mfma(T *a, int64_t ld, int32_t c0, int32_t jhi,
                                                       int32_t wb, const Panel *pan,
                                                       const uint32_t *rp, int64_t ldr, Mod md,
                                                       Digits dg) {
    rp_mfma_body<ND, T>(a, ld, c0, jhi, wb, pan, rp, ldr, md, dg, (int32_t)blockIdx.x);
}

// the side stream's first launch of panel j in the loop ordered by device
// words: one thread waits until panel j's k_panel_narrow launch has begun,
// so the select of panel j that the Rp launch after this one reads is
// complete
__global__ void k_side_wait(LoopSync ls, int32_t j) {
    if (threadIdx.x == 0) ls_wait(ls, LS_NAR, ls.s0 + j);
}

// k_panel_rp_mfma for the loop ordered by device words (after
// k_side_wait): the block at column wb (the next panel's, whose pivot rows
// k_side_done writes) has no Rp to compute and instead waits until the
// select of the next panel has begun (w[32] reaching nwant), so that this
// launch ends only after this panel's k_panel_narrow launch (before that
// select on the main stream): its multipliers, which the update after this
// launch reads, are written and its prune's reads of the rows that the
// update writes are done.  Nothing this
// launch writes is read by that launch: the pivot rows' entries outside
// [wb, wb + 64).
template <int ND, typename T>
__global__ __launch_bounds__(256) void k_panel_rp_look(T *a, int64_t ld, int32_t c0, int32_t jhi,
                                                       int32_t wb, const Panel *pan,
                                                       const uint32_t *rp, int64_t ldr, Mod md,
                                                       Digits dg, LoopSync ls, int32_t nwant) {
    if (c0 + (int32_t)blockIdx.x * PB == wb) {
        if (threadIdx.x == 0) ls_wait(ls, LS_SEL, nwant);
        return;
    }
    rp_mfma_body<ND, T>(a, ld, c0, jhi, -1, pan, rp, ldr, md, dg, (int32_t)blockIdx.x);
}

// the end of the side stream's work of panel j: the pivot rows' entries of
// the next panel's columns [c1, c2) from rp (every reader of the old ones,
// panel j's k_panel_narrow launch, is complete), then w[64] and the host's
// copy of it
template <typename T>
__global__ __launch_bounds__(256) void k_side_done(T *a, int64_t ld, int32_t c0, int32_t c1,
                                                   int32_t c2, const Panel *pan, const uint32_t *rp,
                                                   int64_t ldr, LoopSync ls, int32_t j) {
    const int kk = pan->k;
    for (int i = threadIdx.x; i < kk * PB; i += 256) {
        const int t = i >> 6, jj = i & 63;
        if (c1 + jj < c2) a[(int64_t)pan->erow[t] * ld + c1 + jj] = (T)rp[(int64_t)t * ldr + (c1 + jj - c0)];
    }
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        ls_post(ls, LS_SIDE, ls.s0 + j);
        if (ls.hprog) __hip_atomic_store(ls.hprog, ls.s0 + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

static int grid_of(int64_t n, int bs) {
    int64_t g = (n + bs - 1) / bs;
    if (g < 1) g = 1;
    if (g > 65535 * 4) g = 65535 * 4;
    return (int)g;
}

// element counts of the scratch arrays of an m x n echelon form, in the
// order of DenseWs; the digit planes are empty in the plain form
constexpr int NWS = 11;  // pieces of DenseWs

static void dense_ws_counts(int32_t m, int32_t n, int form, uint32_t p, size_t cnt[NWS],
                            size_t elt[NWS]) {
    const int64_t ldr = ((int64_t)n + 63) & ~63ll;
    const int nd = form == GBGPU_FORM_MFMA ? digits_for(p) : 0;
    const size_t c[NWS] = {(size_t)m, (size_t)n, (size_t)m, (size_t)m, (size_t)m * PB,
                           (size_t)PB * ldr, (size_t)nd * m * PB, (size_t)nd * ldr * PB, 2, 1,
                           (size_t)m};
    const size_t e[NWS] = {4, 4, 4, 4, 4, 4, 1, 1, sizeof(Panel), sizeof(ListState), 4};
    for (int i = 0; i < NWS; ++i) {
        cnt[i] = c[i];
        elt[i] = e[i];
    }
}

size_t dense_ws_bytes(int32_t m, int32_t n, int form, uint32_t p) {
    size_t cnt[NWS], elt[NWS], tot = 0;
    dense_ws_counts(m, n, form, p, cnt, elt);
    for (int i = 0; i < NWS; ++i) tot += (cnt[i] * elt[i] + 255) & ~(size_t)255;
    return tot;
}

void dense_ws_carve(uint8_t *base, int32_t m, int32_t n, int form, uint32_t p, DenseWs *ws) {
    size_t cnt[NWS], elt[NWS], off = 0;
    dense_ws_counts(m, n, form, p, cnt, elt);
    void *ptr[NWS];
    for (int i = 0; i < NWS; ++i) {
        ptr[i] = cnt[i] ? (void *)(base + off) : nullptr;
        off += (cnt[i] * elt[i] + 255) & ~(size_t)255;
    }
    ws->rowpiv = (int32_t *)ptr[0];
    ws->colpiv = (int32_t *)ptr[1];
    ws->active = (int32_t *)ptr[2];
    ws->cand = (int32_t *)ptr[3];
    ws->L = (uint32_t *)ptr[4];
    ws->rp = (uint32_t *)ptr[5];
    ws->dl = (int8_t *)ptr[6];
    ws->dr = (int8_t *)ptr[7];
    ws->pan = (Panel *)ptr[8];
    ws->st = (ListState *)ptr[9];
    ws->alive = (int32_t *)ptr[10];
}

int dense_rref(gbgpu_dense *d, int form) {
    gbgpu_ctx *ctx = d->ctx;
    dfree(d->orow);
    dfree(d->olead);
    d->rank = 0;
    uint8_t *base = nullptr;
    int rc;
    if ((rc = dalloc(ctx, &base, dense_ws_bytes(d->m, d->n, form, d->p)))) return rc;
    DenseWs ws;
    dense_ws_carve(base, d->m, d->n, form, d->p, &ws);
    rc = dense_rref_ws(d, form, ws);
    if (!rc) rc = dense_output(d, ws.colpiv);
    dfree(base);
    if (rc) return rc;
    float ms = 0;
    if (hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]) == hipSuccess) ctx->phase_ms[1] = ms;
    return GBGPU_OK;
}

// the panel loop on A's entries as T (a = A, ld its row stride in T)
template <typename T>
static int panel_loop(gbgpu_dense *d, T *a, int64_t ld, int form, const DenseWs &ws) {
    gbgpu_ctx *ctx = d->ctx;
    hipStream_t s = ctx->stream;
    const int32_t m = d->m, n = d->n;
    const Mod md = Mod::make(d->p);
    const bool big = d->p >= (1u << 17);
    int32_t *rowpiv = ws.rowpiv, *colpiv = ws.colpiv, *active = ws.active, *cand = ws.cand;
    int32_t *alive = ws.alive;
    uint32_t *L = ws.L, *rp = ws.rp;
    int8_t *dl = ws.dl, *dr = ws.dr;
    Panel *pan = ws.pan;
    ListState *st = ws.st;
    const int64_t ldr = ((int64_t)n + 63) & ~63ll;
    Digits dg;
    if (form == GBGPU_FORM_MFMA) digits_setup(dg, d->p, digits_for(d->p));
    dg.L = dl;
    dg.R = dr;
    dg.lplane = (int64_t)m * PB;
    dg.rplane = ldr * PB;
    // the MFMA form orders the two streams through device words (unless
    // GBGPU_DENSE_SYNC=events asks for events); this loop's sequence words
    // are s0 .. s0 + panels
    LoopSync ls;
    if (form == GBGPU_FORM_MFMA) {
        const char *e = getenv("GBGPU_DENSE_SYNC");
        if (!(e && strcmp(e, "events") == 0)) {
            ls.w = ctx->dsync;
            ls.herr = ctx->derr;
            ls.hprog = ctx->dprog;
            ls.s0 = ctx->sync_seq;
            ctx->sync_seq += n / PB + 4;
            *ctx->herr = 0;
        }
    }
    k_dense_init<<<grid_of(m > n ? m : n, 256), 256, 0, s>>>(m, n, rowpiv, colpiv, cand, pan, st);
    const uint32_t r64 = (uint32_t)((((unsigned __int128)1) << 64) % d->p);
    // A[r][cb, ce) -= L[r] * Rp for the rows of the current panel's update
    // (the MFMA form may skip a gap [gb, gb + gsz) inside [cb, ce), gb - cb
    // a multiple of 64: one launch for the columns on both sides of it)
    auto update = [&](hipStream_t s, const Panel *pp, int32_t c0, int32_t cb, int32_t ce,
                      int32_t gb = INT_MAX, int32_t gsz = 0) {
        if (cb >= ce) return;
        if (form == GBGPU_FORM_MFMA) {
            // one and two digits: 32 x 64 block tiles (16 x 32 per wave:
            // fewer registers, more waves in flight); three and four: 64 x 64
            const int rows = dg.nd <= 2 ? 32 : 64;
            dim3 g((unsigned)((ce - cb - gsz + 63) / 64), (unsigned)((m + rows - 1) / rows));
            switch (dg.nd) {
                case 1: k_update_mfma<1, 1, 1, 2, T><<<g, 256, 0, s>>>(a, ld, m, cb, ce, c0, pp, active, md, dg, gb, gsz); break;
                case 2: k_update_mfma<2, 1, 1, 2, T><<<g, 256, 0, s>>>(a, ld, m, cb, ce, c0, pp, active, md, dg, gb, gsz); break;
                default:
                    if constexpr (sizeof(T) == 4) {
                        if (dg.nd == 3) k_update_mfma<3><<<g, 256, 0, s>>>(a, ld, m, cb, ce, c0, pp, active, md, dg, gb, gsz);
                        else k_update_mfma<4><<<g, 256, 0, s>>>(a, ld, m, cb, ce, c0, pp, active, md, dg, gb, gsz);
                    }
                    break;
            }
        } else if constexpr (sizeof(T) == 4) {
            dim3 g((unsigned)((ce - cb + 255) / 256), (unsigned)((m + RT - 1) / RT));
            if (big) k_update_plain<true><<<g, 256, 0, s>>>(a, ld, m, cb, ce, c0, pp, rp, ldr, L, active, md, r64);
            else k_update_plain<false><<<g, 256, 0, s>>>(a, ld, m, cb, ce, c0, pp, rp, ldr, L, active, md, r64);
        }
    };
    if (hipEventRecord(ctx->ev[2], s) != hipSuccess) { return check_hip(ctx, hipGetLastError(), "event"); }
    // Look-ahead: once the next panel's columns are updated, its select (one
    // workgroup) runs while the rest of this panel's Rp and update fill the
    // other CUs from the side stream.  The chain select - prune - Rp of the
    // next panel's columns - gather - update of those columns - select stays
    // on the main stream, which waits for the side stream before the prune
    // (the side stream's work is the longer one in the first panels, where
    // it covers most of the columns).  Panel records alternate
    // between two slots; the candidate list is touched by the select and the
    // prune only, both on the main stream.
    hipStream_t x = ctx->side;
    auto select = [&](hipStream_t sx, int32_t cs, Panel *pp, int32_t j) {
        if (d->p < (1u << 16))
            k_panel_select<1, T><<<1, SEL_T, 0, sx>>>(a, ld, n, cs, cand, rowpiv, colpiv, pp, st, md, dg, ctx->ddone, ls, j);
        else
            k_panel_select<0, T><<<1, SEL_T, 0, sx>>>(a, ld, n, cs, cand, rowpiv, colpiv, pp, st, md, dg, ctx->ddone, ls, j);
    };
    // the host stays at most LAG panels ahead: with the device words it
    // reads the side stream's last panel done from mapped memory
    auto wait_side = [&](int32_t j) -> int {
        const auto t0 = std::chrono::steady_clock::now();
        for (uint32_t i = 0; (int32_t)((uint32_t)*ctx->hprog - (uint32_t)(ls.s0 + j)) < 0; ++i) {
            if ((i & 4095u) == 4095u) {
                const hipError_t e = hipStreamQuery(x);
                if (e != hipSuccess && e != hipErrorNotReady) return check_hip(ctx, e, "dense panel lag");
                if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(60))
                    return set_error(ctx, GBGPU_EDEVICE, "dense panel loop: the side stream made no progress for 60 s");
            }
            std::this_thread::yield();
        }
        return GBGPU_OK;
    };
    // a select that finds no candidate left sets the flag; every entry point
    // has synchronised its stream before it returned, so no earlier loop's
    // select writes it after this
    *ctx->hdone = 0;
    if (m > 0 && n > 0)
        select(s, 0, pan, 0);
    bool side_queued = false;
    constexpr int LAG = (int)(sizeof(ctx->lag_ev) / sizeof(ctx->lag_ev[0]));
    // most panels prune nothing (every block returns at once): the prune's
    // grid stops at four blocks per CU
    const int pg = grid_of(m, 4) < 4 * ctx->num_cu ? grid_of(m, 4) : 4 * ctx->num_cu;
    bool prune_pending = false;
    int it = 0;
    for (int32_t c0 = 0; c0 < n && m > 0; c0 += PB, ++it) {
        if (it >= LAG) {
            // wait for the side stream's work of the panel LAG back (the
            // device still has the ones after it queued; every panel but the
            // last records it); stop once a select found nothing left: the
            // panels from there on are empty
            if (ls.w) {
                const int rc = wait_side(it - LAG);
                if (rc) return rc;
            } else if (hipEventSynchronize(ctx->lag_ev[it % LAG]) != hipSuccess) {
                return check_hip(ctx, hipGetLastError(), "dense panel lag");
            }
            if (*ctx->hdone) break;
        }
        Panel *cur = pan + (it & 1), *nxt = pan + ((it + 1) & 1);
        const int bw = (n - c0) < PB ? (n - c0) : PB;
        const int32_t c1 = c0 + bw;                       // next panel
        const int32_t c2 = (n - c1) < PB ? n : c1 + PB;
        // the MFMA form: the next panel's columns (their Rp, the multipliers,
        // their update) in one launch, the panel's own columns' Rp and the
        // pivot rows' write-back on the side stream; the plain form: Rp over
        // the columns the next select needs first, the rest after that
        // select is on its way
        const bool fused = form == GBGPU_FORM_MFMA && c1 < n;
        if (fused) {
            // with the last panel's prune, if one is pending, in extra blocks
            const int32_t gn = (m + PB - 1) / PB;
            const unsigned g = (unsigned)(gn + (prune_pending ? pg : 0));
            switch (dg.nd) {
                case 1: k_panel_narrow<1, T><<<g, NAR_T, 0, s>>>(a, ld, m, n, c0, c1, c2, rowpiv, cur, active, rp, ldr, md, dg, gn, cand, st, alive, ls, it); break;
                case 2: k_panel_narrow<2, T><<<g, NAR_T, 0, s>>>(a, ld, m, n, c0, c1, c2, rowpiv, cur, active, rp, ldr, md, dg, gn, cand, st, alive, ls, it); break;
                default:
                    if constexpr (sizeof(T) == 4) {
                        if (dg.nd == 3) k_panel_narrow<3, T><<<g, NAR_T, 0, s>>>(a, ld, m, n, c0, c1, c2, rowpiv, cur, active, rp, ldr, md, dg, gn, cand, st, alive, ls, it);
                        else k_panel_narrow<4, T><<<g, NAR_T, 0, s>>>(a, ld, m, n, c0, c1, c2, rowpiv, cur, active, rp, ldr, md, dg, gn, cand, st, alive, ls, it);
                    }
                    break;
            }
            prune_pending = false;
        } else {
            launch_rp<T>(s, true, a, ld, c0, c0, c2, cur, rp, ldr, md, r64, dg);
            k_panel_gather<T><<<grid_of(m, 4), 256, 0, s>>>(a, ld, m, c0, bw, rowpiv, cur, L, active,
                                                             d->p, dg);
        }
        if (c1 < n) {
            if (!fused) update(s, cur, c0, c1, c2);
            if (!ls.w && hipEventRecord(ctx->sync_ev[0], s) != hipSuccess) {
                return check_hip(ctx, hipGetLastError(), "dense look-ahead");
            }
            // the select reads the columns [c1, c2) of candidate rows, which
            // the side stream's work does not touch; queued first, it takes
            // its CU before the side stream's blocks fill the device
            select(s, c1, nxt, it + 1);
            if (!ls.w && hipStreamWaitEvent(x, ctx->sync_ev[0], 0) != hipSuccess) {
                return check_hip(ctx, hipGetLastError(), "dense look-ahead");
            }
            if (fused && ls.w) {
                // Rp of [c0, c1) and [c2, n) once the select of this panel is
                // complete; the launch ends after this panel's
                // k_panel_narrow launch (its multipliers)
                k_side_wait<<<1, 64, 0, x>>>(ls, it);
                const unsigned g = (unsigned)((n - c0 + PB - 1) / PB);
                const int32_t nwant = ls.s0 + it + 1;
                switch (dg.nd) {
                    case 1: k_panel_rp_look<1, T><<<g, 256, 0, x>>>(a, ld, c0, n, c1, cur, rp, ldr, md, dg, ls, nwant); break;
                    case 2: k_panel_rp_look<2, T><<<g, 256, 0, x>>>(a, ld, c0, n, c1, cur, rp, ldr, md, dg, ls, nwant); break;
                    default:
                        if constexpr (sizeof(T) == 4) {
                            if (dg.nd == 3) k_panel_rp_look<3, T><<<g, 256, 0, x>>>(a, ld, c0, n, c1, cur, rp, ldr, md, dg, ls, nwant);
                            else k_panel_rp_look<4, T><<<g, 256, 0, x>>>(a, ld, c0, n, c1, cur, rp, ldr, md, dg, ls, nwant);
                        }
                        break;
                }
            } else if (fused) {
                // Rp of [c0, c1) and [c2, n) on the matrix cores, the pivot
                // rows of [c1, c2) from rp
                const unsigned g = (unsigned)((n - c0 + PB - 1) / PB);
                switch (dg.nd) {
                    case 1: k_panel_rp_mfma<1, T><<<g, 256, 0, x>>>(a, ld, c0, n, c1, cur, rp, ldr, md, dg); break;
                    case 2: k_panel_rp_mfma<2, T><<<g, 256, 0, x>>>(a, ld, c0, n, c1, cur, rp, ldr, md, dg); break;
                    default:
                        if constexpr (sizeof(T) == 4) {
                            if (dg.nd == 3) k_panel_rp_mfma<3, T><<<g, 256, 0, x>>>(a, ld, c0, n, c1, cur, rp, ldr, md, dg);
                            else k_panel_rp_mfma<4, T><<<g, 256, 0, x>>>(a, ld, c0, n, c1, cur, rp, ldr, md, dg);
                        }
                        break;
                }
            } else if (c2 < n) {
                launch_rp<T>(x, false, a, ld, c0, c2, n, cur, rp, ldr, md, r64, dg);
            }
            // one and two digits: 32 x 64 tiles below 4096 rows, the strip
            // walk from there on (la_option 1's 8000 x 18000 dense part:
            // 27.1 -> 23.0 ms; the 4096 x 8192 block 1-2 % faster)
            if (form == GBGPU_FORM_MFMA && c2 < n && c1 - c0 == PB && dg.nd <= 2 && m < 4096) {
                update(x, cur, c0, c0, n, c1, c2 - c1);  // [c0, c1) and [c2, n)
            } else if (form == GBGPU_FORM_MFMA && c2 < n && c1 - c0 == PB) {
                // [c0, c1) and [c2, n) for three and four digits and for tall
                // blocks: 64-row blocks that walk column strips with the Rp
                // digits staged in LDS (measured faster than 64 x 64 tiles
                // for three and four digits; 128-row blocks measured 6 %
                // slower on la_option 1's dense part)
                const int rows = 64;
                // 256-column strips (8 tiles): 512 measured 2 % slower for
                // two digits on la_option 1's dense part and 4-7 % slower for
                // three and four on the 4096 x 8192 block, 1024 8 % slower
                const int sw = 256;
                dim3 g((unsigned)((n - c0 - (c2 - c1) + sw - 1) / sw), (unsigned)((m + rows - 1) / rows));
                switch (dg.nd) {
                    case 1: k_update_wide<1, 1, T, 1, 8><<<g, 256, 0, x>>>(a, ld, m, c0, n, c0, cur, active, md, dg, c1, c2 - c1); break;
                    case 2: k_update_wide<2, 1, T, 1, 8><<<g, 256, 0, x>>>(a, ld, m, c0, n, c0, cur, active, md, dg, c1, c2 - c1); break;
                    default:
                        if constexpr (sizeof(T) == 4) {
                            if (dg.nd == 3) k_update_wide<3, 1, T, 1, 8><<<g, 256, 0, x>>>(a, ld, m, c0, n, c0, cur, active, md, dg, c1, c2 - c1);
                            else k_update_wide<4, 1, T, 1, 8><<<g, 256, 0, x>>>(a, ld, m, c0, n, c0, cur, active, md, dg, c1, c2 - c1);
                        }
                        break;
                }
            } else {
                update(x, cur, c0, c0, c1);
                update(x, cur, c0, c2, n);
            }
            if (ls.w) {
                // the pivot rows of [c1, c2), then the words that release the
                // main stream's wait (queued after this) and the host
                k_side_done<T><<<1, 256, 0, x>>>(a, ld, c0, c1, c2, cur, rp, ldr, ls, it);
                side_queued = true;
                k_main_wait<<<1, 64, 0, s>>>(ls, it);
            } else {
                // the event that orders the side stream's work before the
                // prune is also the one the host waits on LAG panels later
                if (hipEventRecord(ctx->lag_ev[it % LAG], x) != hipSuccess) {
                    return check_hip(ctx, hipGetLastError(), "dense look-ahead");
                }
                if (hipStreamWaitEvent(s, ctx->lag_ev[it % LAG], 0) != hipSuccess) {
                    return check_hip(ctx, hipGetLastError(), "dense look-ahead");
                }
            }
            // the prune: in the MFMA form inside the next panel's
            // k_panel_narrow (nothing reads the list before the select after
            // it), else its own launch
            if (form == GBGPU_FORM_MFMA) prune_pending = true;
            else k_prune<T><<<pg, 256, 0, s>>>(a, ld, n, c1, cand, st, alive);
        } else {
            update(s, cur, c0, c0, n);
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) { return check_hip(ctx, e, "dense panel launch"); }
    }
    // the side stream's last kernel has posted its word but may not have
    // ended: the main stream waits for it once
    if (side_queued) {
        if (hipEventRecord(ctx->lag_ev[0], x) != hipSuccess || hipStreamWaitEvent(s, ctx->lag_ev[0], 0) != hipSuccess)
            return check_hip(ctx, hipGetLastError(), "dense look-ahead");
    }
    (void)hipEventRecord(ctx->ev[3], s);
    return GBGPU_OK;
}

// In-place change of an m x n block's entries between uint32_t and uint16_t
// (residues of p <= 65536): row r's 16-bit entries take the first half of
// its own bytes (row stride 2 ld in 16-bit units).  One wave per row, in
// groups of 512 entries: a group's 16-bit image covers the 32-bit entries of
// half as many groups (to 16 bits) or its 32-bit image twice as many (back),
// so going up (down) the row, every entry a group overwrites has been read
// already, by this wave.
__global__ void k_to_u16(uint32_t *a, int64_t ld, int32_t m, int32_t n) {
    const int lane = threadIdx.x & 63;
    const int64_t w = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (w >= m) return;
    uint32_t *row = a + w * ld;
    uint16_t *out = (uint16_t *)row;
    for (int32_t g = 0; g < n; g += 512) {
        uint32_t v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int32_t j = g + 64 * k + lane;
            v[k] = j < n ? row[j] : 0u;
        }
        __builtin_amdgcn_s_waitcnt(0);  // the group's reads are in before its writes
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int32_t j = g + 64 * k + lane;
            if (j < n) out[j] = (uint16_t)v[k];
        }
    }
}

__global__ void k_to_u32(uint32_t *a, int64_t ld, int32_t m, int32_t n) {
    const int lane = threadIdx.x & 63;
    const int64_t w = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (w >= m) return;
    uint32_t *row = a + w * ld;
    const uint16_t *in = (const uint16_t *)row;
    for (int32_t g = ((n - 1) / 512) * 512; g >= 0; g -= 512) {
        uint32_t v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int32_t j = g + 64 * k + lane;
            v[k] = j < n ? (uint32_t)in[j] : 0u;
        }
        __builtin_amdgcn_s_waitcnt(0);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int32_t j = g + 64 * k + lane;
            if (j < n) row[j] = v[k];
        }
    }
}

int loop_sync_check(gbgpu_ctx *ctx) {
    if (!*ctx->herr) return GBGPU_OK;
    *ctx->herr = 0;
    return set_error(ctx, GBGPU_EDEVICE,
                     "dense panel loop: a wait between its two streams ran out of its bound "
                     "(kernels run one at a time? GBGPU_DENSE_SYNC=events orders them by events)");
}

int dense_rref_ws(gbgpu_dense *d, int form, const DenseWs &ws) {
    // the MFMA form keeps residues of p <= 65521 in 16 bits while it runs,
    // which halves the bytes every update moves
    const bool h = form == GBGPU_FORM_MFMA && d->p <= 65521u && d->m > 0 && d->n > 0;
    if (!h) return panel_loop<uint32_t>(d, d->a, d->ld, form, ws);
    gbgpu_ctx *ctx = d->ctx;
    k_to_u16<<<grid_of(d->m, 4), 256, 0, ctx->stream>>>(d->a, d->ld, d->m, d->n);
    const int rc = panel_loop<uint16_t>(d, (uint16_t *)d->a, 2 * d->ld, form, ws);
    k_to_u32<<<grid_of(d->m, 4), 256, 0, ctx->stream>>>(d->a, d->ld, d->m, d->n);
    if (rc) return rc;
    GBGPU_HIP(ctx, hipGetLastError());
    return GBGPU_OK;
}

int dense_output(gbgpu_dense *d, const int32_t *colpiv) {
    gbgpu_ctx *ctx = d->ctx;
    hipStream_t s = ctx->stream;
    const int32_t n = d->n;
    int rc;
    dfree(d->orow);
    dfree(d->olead);
    d->rank = 0;
    int32_t *flag = nullptr, *rank = nullptr;
    if ((rc = dalloc(ctx, &flag, (size_t)n)) || (rc = dalloc(ctx, &rank, (size_t)n))) {
        dfree(flag);
        return rc;
    }
    if (n > 0) k_out_flags<<<grid_of(n, 256), 256, 0, s>>>(n, colpiv, flag);
    // the lists sized for a full rank, so that the rank is read in the one
    // wait for the stream, with the lists written
    if ((rc = exclusive_scan_i32(ctx, flag, rank, n, nullptr)) ||
        (rc = dalloc(ctx, &d->orow, (size_t)n)) || (rc = dalloc(ctx, &d->olead, (size_t)n))) {
        (void)hipStreamSynchronize(s);
        dfree(flag); dfree(rank); return rc;
    }
    if (n > 0) {
        k_out_list<<<grid_of(n, 256), 256, 0, s>>>(n, colpiv, flag, rank, d->orow, d->olead);
        k_scan_total<int32_t><<<1, 1, 0, s>>>(flag, rank, n, nullptr, ctx->dscan);
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    flush_later(ctx);
    dfree(flag);
    dfree(rank);
    if (e != hipSuccess) return check_hip(ctx, e, "dense output");
    if ((rc = loop_sync_check(ctx))) return rc;
    d->rank = n > 0 ? (int32_t)ctx->hscan[0] : 0;
    return GBGPU_OK;
}

// ------------------------------------------------------------ batches
//
// The panel loop over many dense blocks at once: step k works on columns
// [64 k, 64 k + 64) of every block that is that wide.  Block counts per
// (step, kernel, matrix) depend on the shapes only; the host precomputes
// their prefix sums (tab[step][kind][0..nmat]) and every batch kernel finds
// its matrix by a binary search over them.  No host synchronisation is
// needed inside the loop, so the whole sequence can be captured in a graph.

__device__ __forceinline__ int tab_find(const int32_t *pre, int32_t nmat, int32_t b) {
    int lo = 0, hi = nmat - 1;  // last i with pre[i] <= b
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pre[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return __builtin_amdgcn_readfirstlane(lo);
}

// one block per matrix: rowpiv = colpiv = -1, the candidate list = the rows
// flagged nonzero (all rows without flags), panel state reset
__global__ __launch_bounds__(256) void k_batch_dense_init(const DenseDesc *D) {
    __shared__ int32_t wsum[4];
    const DenseDesc &d = D[blockIdx.x];
    const int tid = threadIdx.x, wv = tid >> 6;
    const int32_t m = d.m, n = d.n;
    for (int32_t i = tid; i < m; i += 256) d.rowpiv[i] = -1;
    for (int32_t i = tid; i < n; i += 256) d.colpiv[i] = -1;
    int32_t out = 0;
    for (int32_t base = 0; base < m; base += 256) {
        const int32_t i = base + tid;
        const int f = i < m ? (d.dnz ? d.dnz[i] : 1) : 0;
        const int incl = wave_incl_scan(f);
        wsum[wv] = __shfl(incl, 63, 64);  // every lane stores the wave total
        __syncthreads();
        int off = 0, tot = 0;
        for (int q = 0; q < 4; ++q) {
            const int ws = wsum[q];
            if (q < wv) off += ws;
            tot += ws;
        }
        if (f) d.cand[out + off + incl - 1] = i;
        out += tot;
        __syncthreads();
    }
    if (tid == 0) {
        d.pan->k = 0;
        d.st->done = out == 0 ? 1 : 0;
        d.st->head = 0;
        d.st->count = out;
        d.st->nzero = 0;
    }
}

__global__ __launch_bounds__(SEL_T) void k_batch_select(const DenseDesc *D, int32_t c0) {
    __shared__ SelShared sh;
    const DenseDesc &d = D[blockIdx.x];
    if (c0 >= d.n) return;
    select_body<2>(sh, d.a, d.ld, d.n, c0, d.cand, d.rowpiv, d.colpiv, d.pan, d.st, d.md, d.dg,
                   nullptr);
}

// the MFMA form's Rp of every matrix: a block per 64 columns, on the matrix
// cores (the plain Rp only the plain form's update reads is not written)
template <int ND>
__global__ __launch_bounds__(256) void k_batch_rp_mfma(const DenseDesc *D, const int32_t *pre,
                                                       int32_t nmat, int32_t c0) {
    const int i = tab_find(pre, nmat, blockIdx.x);
    const DenseDesc &d = D[i];
    rp_mfma_body<ND, uint32_t>(d.a, d.ld, c0, d.n, INT_MIN, d.pan, d.rp, d.ldr, d.md, d.dg,
                               (int32_t)blockIdx.x - pre[i]);
}

__global__ __launch_bounds__(256) void k_batch_rp(const DenseDesc *D, const int32_t *pre,
                                                  int32_t nmat, int32_t c0) {
    const int i = tab_find(pre, nmat, blockIdx.x);
    const DenseDesc &d = D[i];
    rp_body(d.a, d.ld, d.n, c0, d.pan, d.rp, d.ldr, d.md, d.r64, blockIdx.x - pre[i], d.dg);
}

__global__ void k_batch_gather(const DenseDesc *D, const int32_t *pre, int32_t nmat, int32_t c0) {
    const int i = tab_find(pre, nmat, blockIdx.x);
    const DenseDesc &d = D[i];
    const int32_t nb = pre[i + 1] - pre[i];
    const int32_t bw = min(PB, d.n - c0);
    gather_body(d.a, d.ld, d.m, c0, bw, d.rowpiv, d.pan, d.L, d.active,
                (blockIdx.x - pre[i]) * 4 + (threadIdx.x >> 6), nb * 4, d.md.p, d.dg);
}

template <int ND>
__global__ __launch_bounds__(256) void k_batch_update_mfma(const DenseDesc *D, const int32_t *pre,
                                                           int32_t nmat, int32_t c0) {
    const int i = tab_find(pre, nmat, blockIdx.x);
    const DenseDesc &d = D[i];
    const int32_t l = blockIdx.x - pre[i];
    const int32_t ct = (d.n - c0 + 63) / 64;
    update_i8_body<ND, 1, 2, 2>(d.a, d.ld, d.m, c0, d.n, c0, d.pan, d.active, d.md, d.dg, l % ct, l / ct);
}

// blocks that walk strips of 32 NT columns of 64 RS rows with the Rp digits
// staged in LDS (K = 64), as the single block's side stream does
template <int ND, int RS, int NT>
__global__ __launch_bounds__(256) void k_batch_update_wide(const DenseDesc *D, const int32_t *pre,
                                                           int32_t nmat, int32_t c0) {
    constexpr int W = NT * 16 * WT_CS;
    const int i = tab_find(pre, nmat, blockIdx.x);
    const DenseDesc &d = D[i];
    const int32_t l = blockIdx.x - pre[i];
    const int32_t ct = (d.n - c0 + W - 1) / W;
    update_wide_body<ND, RS, uint32_t, 1, NT>(d.a, d.ld, d.m, c0, d.n, c0, d.pan, d.active, d.md,
                                              d.dg, INT_MAX, 0, l % ct, l / ct);
}

// the MFMA form's batch update kernel for nd digits and tile shape (cols x
// rows per block: 64 x 64 tiles, or strips of 256 or 512 columns of 64 or,
// for nd <= 2, 128 rows); nullptr for a shape without a kernel
static const void *batch_update_fn(int nd, int cols, int rows) {
#define GBGPU_BU(ND)                                                                        \
    if (cols == 64 && rows == 64) return (const void *)k_batch_update_mfma<ND>;             \
    if (cols == 256 && rows == 64) return (const void *)k_batch_update_wide<ND, 1, 8>;      \
    if (cols == 512 && rows == 64) return (const void *)k_batch_update_wide<ND, 1, 16>;     \
    if (ND <= 2 && cols == 256 && rows == 128) return (const void *)k_batch_update_wide<(ND <= 2 ? ND : 2), 2, 8>;  \
    if (ND <= 2 && cols == 512 && rows == 128) return (const void *)k_batch_update_wide<(ND <= 2 ? ND : 2), 2, 16>; \
    return nullptr;
    switch (nd) {
        case 1: { GBGPU_BU(1) }
        case 2: { GBGPU_BU(2) }
        case 3: { GBGPU_BU(3) }
        default: { GBGPU_BU(4) }
    }
#undef GBGPU_BU
}

bool batch_update_shape_ok(int nd, int cols, int rows) {
    return batch_update_fn(nd, cols, rows) != nullptr;
}

template <bool BIG>
__global__ __launch_bounds__(256) void k_batch_update_plain(const DenseDesc *D, const int32_t *pre,
                                                            int32_t nmat, int32_t c0) {
    const int i = tab_find(pre, nmat, blockIdx.x);
    const DenseDesc &d = D[i];
    const int32_t l = blockIdx.x - pre[i];
    const int32_t ct = (d.n - c0 + 255) / 256;
    update_plain_body<BIG>(d.a, d.ld, d.m, c0, d.n, c0, d.pan, d.rp, d.ldr, d.L, d.active, d.md,
                           d.r64, l % ct, l / ct);
}

__global__ void k_batch_prune_flags(const DenseDesc *D, const int32_t *pre, int32_t nmat,
                                    int32_t c0) {
    const int i = tab_find(pre, nmat, blockIdx.x);
    const DenseDesc &d = D[i];
    const int32_t nb = pre[i + 1] - pre[i];
    prune_flags_body(d.a, d.ld, d.n, c0 + PB, d.cand, d.st, d.active,
                     (blockIdx.x - pre[i]) * 4 + (threadIdx.x >> 6), nb * 4);
}

__global__ __launch_bounds__(256) void k_batch_prune_compact(const DenseDesc *D, int32_t c0) {
    const DenseDesc &d = D[blockIdx.x];
    if (c0 + PB >= d.n) return;
    prune_compact_body(d.cand, d.st, d.active);
}

size_t dense_panel_bytes() { return sizeof(Panel); }
size_t dense_liststate_bytes() { return sizeof(ListState); }

int batch_dense_init(gbgpu_ctx *ctx, const DenseDesc *D, int32_t nmat) {
    int rc;
    if (nmat > 0 && (rc = check_launch(ctx, (const void *)k_batch_dense_init, nmat, 256, "batch init")))
        return rc;
    if (nmat > 0) k_batch_dense_init<<<nmat, 256, 0, ctx->stream>>>(D);
    GBGPU_HIP(ctx, hipGetLastError());
    return GBGPU_OK;
}

// a launch's shape against the device's limits, checked on the host before
// it is issued (or captured): a nonzero grid, a block the kernel was
// compiled for, grid x block within the 32-bit dispatch size, the kernel's
// static LDS within the per-workgroup limit
int check_launch(gbgpu_ctx *ctx, const void *fn, int64_t grid, int block, const char *what) {
    static std::mutex mu;
    static std::unordered_map<const void *, hipFuncAttributes> seen;
    static int lds_max = 0;
    hipFuncAttributes at;
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = seen.find(fn);
        if (it == seen.end()) {
            GBGPU_HIP(ctx, hipFuncGetAttributes(&at, fn));
            seen[fn] = at;
            if (!lds_max)
                GBGPU_HIP(ctx, hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock,
                                                     ctx->device));
        } else {
            at = it->second;
        }
    }
    if (grid < 1 || block < 1 || block > at.maxThreadsPerBlock || grid * (int64_t)block >= (1ll << 32) ||
        (int64_t)at.sharedSizeBytes > (int64_t)lds_max)
        return set_error(ctx, GBGPU_EDEVICE,
                         std::string("launch outside the device limits: ") + what + " grid " +
                             std::to_string(grid) + " block " + std::to_string(block) + " lds " +
                             std::to_string(at.sharedSizeBytes));
    return GBGPU_OK;
}

int batch_dense_step(gbgpu_ctx *ctx, const DenseDesc *D, int32_t nmat, const int32_t *tab,
                     const int32_t *htab, int32_t step, int form, bool big, int nd, int ucols,
                     int urows) {
    hipStream_t s = ctx->stream;
    const int32_t c0 = step * PB;
    const int32_t *pre = tab + (int64_t)step * 4 * (nmat + 1);
    const int32_t *hpre = htab + (int64_t)step * 4 * (nmat + 1);
    const int32_t nrp = hpre[nmat], nga = hpre[(nmat + 1) + nmat];
    const int32_t nup = hpre[2 * (nmat + 1) + nmat], npr = hpre[3 * (nmat + 1) + nmat];
    int rc;
    const void *upd = form == GBGPU_FORM_MFMA
                          ? batch_update_fn(nd, ucols, urows)
                          : (big ? (const void *)k_batch_update_plain<true>
                                 : (const void *)k_batch_update_plain<false>);
    if (!upd) return set_error(ctx, GBGPU_EINVAL, "no batch update kernel of that shape");
    const void *rpk = form != GBGPU_FORM_MFMA ? (const void *)k_batch_rp
                      : nd == 1 ? (const void *)k_batch_rp_mfma<1>
                      : nd == 2 ? (const void *)k_batch_rp_mfma<2>
                      : nd == 3 ? (const void *)k_batch_rp_mfma<3>
                                : (const void *)k_batch_rp_mfma<4>;
    if ((rc = check_launch(ctx, (const void *)k_batch_select, nmat, SEL_T, "batch select")) ||
        (nrp > 0 && (rc = check_launch(ctx, rpk, nrp, 256, "batch rp"))) ||
        (nga > 0 && (rc = check_launch(ctx, (const void *)k_batch_gather, nga, 256, "batch gather"))) ||
        (nup > 0 && (rc = check_launch(ctx, upd, nup, 256, "batch update"))) ||
        (npr > 0 && (rc = check_launch(ctx, (const void *)k_batch_prune_flags, npr, 256, "batch prune"))) ||
        (npr > 0 && (rc = check_launch(ctx, (const void *)k_batch_prune_compact, nmat, 256, "batch prune"))))
        return rc;
    k_batch_select<<<nmat, SEL_T, 0, s>>>(D, c0);
    if (nrp > 0) {
        int32_t nm = nmat, cc = c0;
        void *args[] = {(void *)&D, (void *)&pre, (void *)&nm, (void *)&cc};
        GBGPU_HIP(ctx, hipLaunchKernel(rpk, dim3((unsigned)nrp), dim3(256), args, 0, s));
    }
    if (nga > 0) k_batch_gather<<<nga, 256, 0, s>>>(D, pre + (nmat + 1), nmat, c0);
    if (nup > 0) {
        const int32_t *pu = pre + 2 * (nmat + 1);
        if (form == GBGPU_FORM_MFMA) {
            int32_t nm = nmat, cc = c0;
            void *args[] = {(void *)&D, (void *)&pu, (void *)&nm, (void *)&cc};
            GBGPU_HIP(ctx, hipLaunchKernel(upd, dim3((unsigned)nup), dim3(256), args, 0, s));
        } else {
            if (big) k_batch_update_plain<true><<<nup, 256, 0, s>>>(D, pu, nmat, c0);
            else k_batch_update_plain<false><<<nup, 256, 0, s>>>(D, pu, nmat, c0);
        }
    }
    if (npr > 0) {
        k_batch_prune_flags<<<npr, 256, 0, s>>>(D, pre + 3 * (nmat + 1), nmat, c0);
        k_batch_prune_compact<<<nmat, 256, 0, s>>>(D, c0);
    }
    GBGPU_HIP(ctx, hipGetLastError());
    return GBGPU_OK;
}

// block counts of one batch step (prefix sums over the matrices)
void batch_dense_counts(const DenseDesc *hd, int32_t nmat, int32_t step, int form, int ucols,
                        int urows, int32_t *out) {
    const int32_t c0 = step * PB;
    int32_t *prp = out, *pga = out + (nmat + 1), *pup = out + 2 * (nmat + 1),
            *ppr = out + 3 * (nmat + 1);
    prp[0] = pga[0] = pup[0] = ppr[0] = 0;
    for (int32_t i = 0; i < nmat; ++i) {
        const int32_t m = hd[i].m, n = hd[i].n;
        const bool on = c0 < n && m > 0;
        const int32_t rows4 = on ? (m + 3) / 4 : 0;
        const int32_t cap = rows4 < 2048 ? rows4 : 2048;
        int32_t nup = 0;
        if (on) {
            if (form == GBGPU_FORM_MFMA) nup = ((n - c0 + ucols - 1) / ucols) * ((m + urows - 1) / urows);
            else nup = ((n - c0 + 255) / 256) * ((m + RT - 1) / RT);
        }
        prp[i + 1] = prp[i] + (on ? (n - c0 + 63) / 64 : 0);
        pga[i + 1] = pga[i] + cap;
        pup[i + 1] = pup[i] + nup;
        ppr[i + 1] = ppr[i] + ((on && c0 + PB < n) ? cap : 0);
    }
}

}  // namespace gbgpu
