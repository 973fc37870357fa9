// Shared pieces of the blocked right-looking factorisations (FP64 and mixed
// precision): the work-group -> block-pair map of the trailing update, the
// per-phase event timer and the one-panel look-ahead schedule.
#pragma once

#include <algorithm>
#include <cstdlib>
#include <utility>
#include <vector>

#include "hs_internal.h"

// Which (I, J) blocks one trailing-update launch of step k covers; the
// launch's work-group q (one per block, times the tiles of a block) decodes
// its pair with pair_of:
//  mode 1: diagonal blocks I = J = rowStart + q             (SYRK entry)
//  mode 2: rows k+2.., J in [k+1, I-1], offset pairOffset    (GEMM entry)
//  mode 3: rows k+1.., J in [k+1, I],   offset pairOffset    (whole step)
//  mode 4: rows k+2.., J in [k+2, I]                         (look-ahead rest)
//  mode 5: J = k+1, I = k+1+q                                (look-ahead panel column)
struct PairMap {
    int k, rowStart, mode;
    i64 pairOffset;
    // band order (modes 3 and 4 over a whole triangle): m rows, bands of
    // `band` block rows walked column by column; m == 0 keeps pair_of's order
    int m = 0, band = 4;
    // XCD chunking: work-group g runs logical tile (g % 8) * per + g / 8, so
    // each XCD (work-groups are dealt to the 8 XCDs round-robin) walks one
    // contiguous stretch of the order; per == 0 keeps the launch order
    i64 total = 0, per = 0;
};

__device__ inline void pair_of(const PairMap& pm, i64 q, int* I, int* J) {
    if (pm.mode == 1) {
        *I = pm.rowStart + (int)q;
        *J = *I;
        return;
    }
    if (pm.mode == 5) {
        *I = pm.k + 1 + (int)q;
        *J = pm.k + 1;
        return;
    }
    int r, c;
    if (pm.mode == 4) {
        hs_tri_decode(q, &r, &c);
        *I = pm.k + 2 + r;
        *J = pm.k + 2 + c;
        return;
    }
    hs_tri_decode(q + pm.pairOffset, &r, &c);
    if (pm.mode == 2) {
        *I = pm.k + 2 + r;
        *J = pm.k + 1 + c;
    } else {
        *I = pm.k + 1 + r;
        *J = pm.k + 1 + c;
    }
}

// Pair q of a triangle of m block rows (relative rows i >= columns j) in
// band order: bands of H rows; inside a band, column j by column j, the
// band's rows below the diagonal. Consecutive work-groups then share the
// B operand L_Jk while the band's H A operands stay in the XCD's L2, so an
// operand block is fetched once per band instead of once per row.
__device__ inline void band_decode(i64 q, int m, int H, int* pi, int* pj) {
    auto S = [&](i64 B) { return (i64)H * H * B * (B - 1) / 2 + B * (i64)H * (H + 1) / 2; };
    int lo = 0, hi = (m + H - 1) / H - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (S(mid) <= q) lo = mid;
        else hi = mid - 1;
    }
    i64 p = q - S(lo);
    const int r0 = lo * H;
    const int hp = min(H, m - r0);
    const i64 rect = (i64)hp * r0;
    if (p < rect) {
        *pj = (int)(p / hp);
        *pi = r0 + (int)(p % hp);
        return;
    }
    p -= rect;
    int t = 0;
    while (p >= hp - t) {
        p -= hp - t;
        ++t;
    }
    *pj = r0 + t;
    *pi = r0 + t + (int)p;
}

// Block pair (I, J) and tile index of this work-group; false for the idle
// work-groups that pad an XCD-chunked launch to a multiple of 8.
__device__ inline bool tile_of(const PairMap& pm, int tilesPerBlock, int* I, int* J, int* tile) {
    i64 g = blockIdx.x;
    if (pm.per > 0) {
        g = (g % 8) * pm.per + g / 8;
        if (g >= pm.total) return false;
    }
    const i64 q = g / tilesPerBlock;
    *tile = (int)(g % tilesPerBlock);
    if (pm.m > 0) {
        int i, j;
        band_decode(q, pm.m, pm.band, &i, &j);
        *I = pm.rowStart + i;
        *J = pm.rowStart + j;
    } else {
        pair_of(pm, q, I, J);
    }
    return true;
}

// Number of block pairs and the pair offset of a launch (see PairMap).
inline void hs_pair_range(int k, int rowStart, int rowCount, int n, int mode, i64* pairs, i64* offset, int* rs) {
    *offset = 0;
    *rs = rowStart;
    if (mode == 1) {
        *pairs = rowCount;
    } else if (mode == 4) {
        const i64 m = n - k - 2;
        *pairs = m > 0 ? m * (m + 1) / 2 : 0;
        *rs = k + 2;
    } else if (mode == 5) {
        *pairs = n - k - 1;
        *rs = k + 1;
    } else if (mode == 2) {
        const int firstRow = std::max(rowStart, k + 2);
        const int lastRow = rowStart + rowCount;
        if (lastRow <= firstRow) {
            *pairs = 0;
            return;
        }
        const i64 r0 = firstRow - (k + 2), r1 = lastRow - (k + 2);
        *pairs = r1 * (r1 + 1) / 2 - r0 * (r0 + 1) / 2;
        *offset = r0 * (r0 + 1) / 2;
    } else {
        const i64 r0 = rowStart - (k + 1), r1 = rowStart + rowCount - (k + 1);
        *pairs = r1 * (r1 + 1) / 2 - r0 * (r0 + 1) / 2;
        *offset = r0 * (r0 + 1) / 2;
    }
}

// Work-group order of a trailing-update launch; HS_TRAIL_ORDER (bit 0 band
// order, bit 1 XCD chunking, default both) and HS_TRAIL_BAND (band height,
// default 4) select it for experiments.
inline void hs_trail_order(int* bits, int* band) {
    static int b0 = -1, h0 = 4;
    if (b0 < 0) {
        const char* e = std::getenv("HS_TRAIL_ORDER");
        b0 = e ? std::atoi(e) : 3;
        const char* hb = std::getenv("HS_TRAIL_BAND");
        h0 = hb ? std::max(1, std::atoi(hb)) : 4;
    }
    *bits = b0;
    *band = h0;
}

// Fills the pair map of a trailing-update launch with tilesPerBlock work-
// groups per block pair; returns the grid size (0: nothing to do).
inline i64 hs_pair_map(int k, int rowStart, int rowCount, int n, int mode, i64 tilesPerBlock, PairMap* pm) {
    i64 pairs, off;
    int rs;
    hs_pair_range(k, rowStart, rowCount, n, mode, &pairs, &off, &rs);
    *pm = PairMap{k, rs, mode, off};
    if (pairs <= 0) return 0;
    int bits, band;
    hs_trail_order(&bits, &band);
    if ((bits & 1) && (mode == 4 || (mode == 3 && off == 0))) {
        pm->m = (mode == 4) ? n - k - 2 : rowCount;
        pm->band = band;
    }
    const i64 total = pairs * tilesPerBlock;
    if ((bits & 2) && total >= 64) {
        pm->total = total;
        pm->per = (total + 7) / 8;
        return 8 * pm->per;
    }
    return total;
}

// Runs launches that use h->stream on another stream of the same handle.
struct OnStream {
    hs_handle_t h;
    hipStream_t saved;
    OnStream(hs_handle_t hh, hipStream_t s) : h(hh), saved(hh->stream) { h->stream = s; }
    ~OnStream() { h->stream = saved; }
};

// Per-phase event bracketing (0 = POTRF, 1 = TRSM, 2 = main-stream trailing
// update), summed after the final synchronisation.
struct PhaseTimer {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> spans[3];
    bool on = false;
    int begin(hipStream_t s, int phase) {
        if (!on) return HS_OK;
        hipEvent_t e;
        HS_HIP(hipEventCreate(&e));
        HS_HIP(hipEventRecord(e, s));
        spans[phase].push_back({e, nullptr});
        return HS_OK;
    }
    int end(hipStream_t s, int phase) {
        if (!on) return HS_OK;
        hipEvent_t e;
        HS_HIP(hipEventCreate(&e));
        HS_HIP(hipEventRecord(e, s));
        spans[phase].back().second = e;
        return HS_OK;
    }
    int collect(double* ms, int64_t* launches) {
        for (int p = 0; p < 3; ++p) {
            double t = 0;
            for (auto& pr : spans[p]) {
                float v = 0;
                HS_HIP(hipEventElapsedTime(&v, pr.first, pr.second));
                t += v;
            }
            ms[p] = t;
            launches[p] = (int64_t)spans[p].size();
        }
        return HS_OK;
    }
    ~PhaseTimer() {
        for (auto& sp : spans)
            for (auto& pr : sp) {
                if (pr.first) hipEventDestroy(pr.first);
                if (pr.second) hipEventDestroy(pr.second);
            }
    }
};

// Right-looking factorisation with a look-ahead of one panel on two streams
// of the same priority. Stream S2 factors panels: it applies step k's update
// to block column k+1 only (trail mode 5), then POTRF and TRSM of panel k+1.
// Stream S1 applies step k's update to the rest of the trailing triangle
// (mode 4, J >= k+2) as soon as panel k is final, so the panel work of k+1
// runs underneath it. S1's update k waits for panel k (event P[k]); S2's
// column update of step k+1 waits for S1's update k, which also wrote block
// column k+2.
//   potrf(k), trsm(k, blockStart, blockCount): launch on h->stream
//   trail(stream, k, mode, rowStart, rowCount): launch on `stream`
template <class Potrf, class Trsm, class Trail>
int hs_lookahead_factor(hs_handle_t h, int n, PhaseTimer& pt, Potrf potrf, Trsm trsm, Trail trail) {
    int rc = hs_ensure_events(h, 2 * (size_t)n + 2);
    if (rc) return rc;
    hipEvent_t* evP = h->events.data();          // panel k final (S2)
    hipEvent_t* evU = h->events.data() + n + 1;  // S1 update k done
    hipStream_t s1 = h->stream, s2 = h->stream2;
    // S2 must see everything already queued on S1 (e.g. the input assembly)
    HS_HIP(hipEventRecord(evU[n], s1));
    HS_HIP(hipStreamWaitEvent(s2, evU[n], 0));
    {
        OnStream os(h, s2);
        if ((rc = pt.begin(s2, 0)) || (rc = potrf(0)) || (rc = pt.end(s2, 0))) return rc;
        if (n > 1 && ((rc = pt.begin(s2, 1)) || (rc = trsm(0, 1, n - 1)) || (rc = pt.end(s2, 1)))) return rc;
    }
    HS_HIP(hipEventRecord(evP[0], s2));
    for (int k = 0; k + 1 < n; ++k) {
        const int m = n - k - 1;
        if (k >= 1) HS_HIP(hipStreamWaitEvent(s2, evU[k - 1], 0));
        if ((rc = trail(s2, k, 5, k + 1, m))) return rc;
        {
            OnStream os(h, s2);
            if ((rc = pt.begin(s2, 0)) || (rc = potrf(k + 1)) || (rc = pt.end(s2, 0))) return rc;
            if (m > 1 && ((rc = pt.begin(s2, 1)) || (rc = trsm(k + 1, k + 2, m - 1)) || (rc = pt.end(s2, 1))))
                return rc;
        }
        HS_HIP(hipEventRecord(evP[k + 1], s2));
        HS_HIP(hipStreamWaitEvent(s1, evP[k], 0));
        if (m > 1) {
            if ((rc = pt.begin(s1, 2)) || (rc = trail(s1, k, 4, k + 2, m - 1)) || (rc = pt.end(s1, 2))) return rc;
        }
        HS_HIP(hipEventRecord(evU[k], s1));
    }
    HS_HIP(hipStreamWaitEvent(s1, evP[n - 1], 0));
    return HS_OK;
}
