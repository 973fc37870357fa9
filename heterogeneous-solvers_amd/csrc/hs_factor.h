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
// its pair with pair_of. A launch applies np consecutive panels k .. k+np-1:
// A_IJ -= sum_p L_I,k+p L_J,k+p^T (np = 2 on the two-panel schedule).
//  mode 1: diagonal blocks I = J = rowStart + q             (SYRK entry)
//  mode 2: rows k+2.., J in [k+1, I-1], offset pairOffset    (GEMM entry)
//  mode 3: rows k+1.., J in [k+1, I],   offset pairOffset    (whole step)
//  mode 4: rows and columns from rs = k + 2 np, J <= I       (look-ahead rest)
//  mode 5: J = rs = k+1, I = rs + q                          (look-ahead panel column)
//  mode 6: J in [rs, rs+nc) with rs = rowStart, nc = rowCount, I >= J
//                                                            (look-ahead columns)
struct PairMap {
    int k, rowStart, mode;
    i64 pairOffset;
    int n = 0, np = 1, nc = 1;
    // band order (modes 3 and 4 over a whole triangle): m rows, bands of
    // `band` block rows walked column by column; m == 0 keeps pair_of's order
    int m = 0, band = 4;
    // XCD chunking: work-group g runs logical tile (g % 8) * per + g / 8, so
    // each XCD (work-groups are dealt to the 8 XCDs round-robin) walks one
    // contiguous stretch of the order; per == 0 keeps the launch order
    i64 total = 0, per = 0;
};

__host__ __device__ inline void pair_of(const PairMap& pm, i64 q, int* I, int* J) {
    if (pm.mode == 1) {
        *I = pm.rowStart + (int)q;
        *J = *I;
        return;
    }
    if (pm.mode == 5) {
        *I = pm.rowStart + (int)q;
        *J = pm.rowStart;
        return;
    }
    if (pm.mode == 6) {
        // columns rs .. rs+nc-1, column c holding rows rs+c .. n-1
        int c = 0;
        i64 r = q;
        while (c + 1 < pm.nc && r >= pm.n - pm.rowStart - c) {
            r -= pm.n - pm.rowStart - c;
            ++c;
        }
        *J = pm.rowStart + c;
        *I = *J + (int)r;
        return;
    }
    int r, c;
    if (pm.mode == 4) {
        hs_tri_decode(q, &r, &c);
        *I = pm.rowStart + r;
        *J = pm.rowStart + c;
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
__host__ __device__ inline void band_decode(i64 q, int m, int H, int* pi, int* pj) {
    auto S = [&](i64 B) { return (i64)H * H * B * (B - 1) / 2 + B * (i64)H * (H + 1) / 2; };
    int lo = 0, hi = (m + H - 1) / H - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (S(mid) <= q) lo = mid;
        else hi = mid - 1;
    }
    i64 p = q - S(lo);
    const int r0 = lo * H;
    const int hp = (m - r0 < H) ? m - r0 : H;
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

// Block pair (I, J) and tile index of work-group g; false for the idle
// work-groups that pad an XCD-chunked launch to a multiple of 8.
__host__ __device__ inline bool tile_of_g(const PairMap& pm, int tilesPerBlock, i64 g, int* I, int* J, int* tile) {
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

__device__ inline bool tile_of(const PairMap& pm, int tilesPerBlock, int* I, int* J, int* tile) {
    return tile_of_g(pm, tilesPerBlock, blockIdx.x, I, J, tile);
}

// Number of block pairs and the pair offset of a launch (see PairMap).
inline void hs_pair_range(int k, int rowStart, int rowCount, int n, int mode, int np, i64* pairs, i64* offset,
                          int* rs) {
    *offset = 0;
    *rs = rowStart;
    if (mode == 1) {
        *pairs = rowCount;
    } else if (mode == 4) {
        *rs = k + 2 * np;
        const i64 m = n - *rs;
        *pairs = m > 0 ? m * (m + 1) / 2 : 0;
    } else if (mode == 5) {
        *rs = k + 1;
        *pairs = std::max(0, n - *rs);
    } else if (mode == 6) {
        *pairs = 0;
        for (int c = 0; c < rowCount; ++c) *pairs += std::max(0, n - rowStart - c);
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

// Work-group order of a trailing-update launch: band order (bit 0) and XCD
// chunking (bit 1). The FP64 update ran equally fast with and without them
// (it is bound by MFMA issue, not by its 1-2.5 TB/s of memory traffic); both
// stay on for the L2 re-reads they save. Band height where a block pair has 4
// or more tiles (128-tile FP64 update on b = 256), from the HBM reads per
// rest-update launch (FETCH_SIZE) at N = 65536, kernel time the same for all:
//   4 panels: 20.7 / 18.6 / 18.4 / 24.8 / 26.9 GB at heights 4 / 2 / 3 / 6 / 8
//             (25.4 GB at height 4 without chunking)
//   6 panels: 45.2 / 32.2 / 35.5 / 41.2 GB at heights 1 / 2 / 3 / 4
// so about 6 MB of a band's operand rows: height 3 up to 4 panels, 2 beyond.
// Height 4 where one tile covers a block (256-tile mixed updates: height 2
// ran 1% slower). HS_TRAIL_BAND = 1..16 and HS_TRAIL_ORDER = 0..3 override.
inline void hs_trail_order(i64 tilesPerBlock, int np, int* bits, int* band) {
    // read once per process (thread-safe static initialisation)
    static const int b = [] {
        const char* e = std::getenv("HS_TRAIL_BAND");
        const int w = e ? std::atoi(e) : 0;
        return (w >= 1 && w <= 16) ? w : 0;
    }();
    static const int o = [] {
        const char* f = std::getenv("HS_TRAIL_ORDER");
        const int v = f ? std::atoi(f) : -1;
        return (v >= 0 && v <= 3) ? v : 3;
    }();
    *bits = o;
    *band = b > 0 ? b : (tilesPerBlock >= 4 ? (np >= 5 ? 2 : 3) : 4);
}

// Fills the pair map of a trailing-update launch with tilesPerBlock work-
// groups per block pair, in the work-group order `bits` with bands of `band`
// block rows; returns the grid size (0: nothing to do).
inline i64 hs_pair_map_order(int k, int rowStart, int rowCount, int n, int mode, int np, i64 tilesPerBlock, int bits,
                             int band, PairMap* pm) {
    i64 pairs, off;
    int rs;
    hs_pair_range(k, rowStart, rowCount, n, mode, np, &pairs, &off, &rs);
    *pm = PairMap{k, rs, mode, off};
    pm->n = n;
    pm->np = np;
    pm->nc = (mode == 6) ? std::max(1, std::min(rowCount, n - rs)) : 1;
    if (pairs <= 0) return 0;
    if ((bits & 1) && (mode == 4 || (mode == 3 && off == 0))) {
        pm->m = (mode == 4) ? n - rs : rowCount;
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

// hs_pair_map_order in the launch order of hs_trail_order
inline i64 hs_pair_map(int k, int rowStart, int rowCount, int n, int mode, int np, i64 tilesPerBlock, PairMap* pm) {
    int bits, band;
    hs_trail_order(tilesPerBlock, np, &bits, &band);
    return hs_pair_map_order(k, rowStart, rowCount, n, mode, np, tilesPerBlock, bits, band, pm);
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
    // phases 0 POTRF, 1 TRSM, 2 rest updates (S1), 3 the look-ahead column
    // updates on the panel stream (timed only when `columns` is set); each span
    // remembers the block column it belongs to (its first panel)
    std::vector<std::pair<hipEvent_t, hipEvent_t>> spans[4];
    std::vector<int> cols[4];
    bool on = false;
    bool columns = false;
    int begin(hipStream_t s, int phase, int col = -1) {
        if (!on || (phase == 3 && !columns)) return HS_OK;
        hipEvent_t e;
        HS_HIP(hipEventCreate(&e));
        HS_HIP(hipEventRecord(e, s));
        spans[phase].push_back({e, nullptr});
        cols[phase].push_back(col);
        return HS_OK;
    }
    int end(hipStream_t s, int phase) {
        if (!on || (phase == 3 && !columns)) return HS_OK;
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
    // per block column k < n: POTRF, TRSM, and every trailing-update launch
    // whose first panel is k (the rest update and the column updates)
    int collect_columns(int n, double* potrf, double* trsm, double* update) {
        double* out[4] = {potrf, trsm, update, update};
        for (int k = 0; k < n; ++k) potrf[k] = trsm[k] = update[k] = 0.0;
        for (int p = 0; p < 4; ++p)
            for (size_t i = 0; i < spans[p].size(); ++i) {
                float v = 0;
                HS_HIP(hipEventElapsedTime(&v, spans[p][i].first, spans[p][i].second));
                const int k = cols[p][i];
                if (k >= 0 && k < n) out[p][k] += v;
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

// Panels per step of the look-ahead schedule: 4 (trailing updates of depth
// 4b). Measured FP64 factor at N = 65536, b = 256: 1569 / 1526 / 1506 ms for
// 2 / 3 / 4 with an earlier update kernel; with the current one 1416-1420 /
// 1408 / 1404-1406 / 1411 ms for 4 / 5 / 6 / 8. Six panels save 0.8% of the
// factor but move work to the panel stream's column updates (the rest update
// then runs at 61.3 instead of 62.4 TFLOP/s) and re-read more operand tiles
// (38.2 instead of 23.6 GB of HBM traffic per launch against 11.2 and 11.3
// algorithmic), so four stay. GP NLL + gradient at N = 16384, b = 128: the
// same for 4 and 6. HS_PANELS = 1..8 chooses another depth, never more than n/2.
inline int hs_schedule_panels(int n) {
    static const int v = [] {
        const char* e = std::getenv("HS_PANELS");
        const int w = e ? std::atoi(e) : 0;
        return (w >= 1 && w <= 8) ? w : 4;
    }();
    return std::max(1, std::min(v, n / 2));
}

// Launches and algorithmic flops of the rest-update launches (stream S1) of
// the schedule for n block columns of size b: step k (k = 0, P, 2P, ...)
// updates the m = n - k - 2P block rows from row k + 2P with P panels, i.e.
// P b^3 m^2 flops (2 P b^3 per off-diagonal block, P b^3 per diagonal block,
// whose upper half is not needed).
inline void hs_rest_update_work(int n, int b, int64_t* launches, double* flops, int panels = 0) {
    const int P = panels > 0 ? std::max(1, std::min(panels, n / 2)) : hs_schedule_panels(n);
    int64_t cnt = 0;
    double f = 0.0;
    for (int k = 0; k + P < n; k += P) {
        const double m = n - k - 2 * P;
        if (m <= 0) continue;
        ++cnt;
        f += (double)P * b * b * (double)b * m * m;
    }
    *launches = cnt;
    *flops = f;
}

// Right-looking factorisation with a look-ahead of one step on two streams of
// the same priority; a step applies P panels (P = hs_schedule_panels(n)).
// Stream S2 factors panels: the next step's columns k+P .. k+2P-1 get the
// update by panels k .. k+P-1 (trail mode 5 for P = 1, mode 6 otherwise),
// then for each of those P panels POTRF and TRSM, each panel updating the
// step's later columns before their POTRF (modes 5 and 6 with one panel). Stream S1 applies the step's panels to
// the rest of the trailing triangle (mode 4, rows and columns >= k + 2P) as
// soon as they are final, so the next step's panel work runs underneath it.
// S1's update of step s waits for the panels of step s (event P[s]); S2's
// column update of step s+1 waits for S1's update of step s-1, which wrote
// those columns. No work between a step's panel chain and the event that
// publishes it.
//
// The schedule is written once, over an executor that issues its operations
// on the device (hs_lookahead_factor) or records them on the host
// (hs_lookahead_dry_run, which the schedule's race analysis replays).
// Streams are 1 (S1, the handle's ordering stream) and 2 (S2, the panel
// stream); events are indices, P[s] = s and U[s] = n + 1 + s (2 n + 2 in all):
//   ex.record(ev, stream), ex.wait(stream, ev)
//   ex.potrf(k), ex.trsm(k, blockStart, blockCount)      on S2
//   ex.trail(stream, k, mode, rowStart, rowCount, np)
//   ex.prep(k0)                                          on S2, after the panel chain of k0
//   ex.begin(stream, phase, k), ex.end(stream, phase)    timing brackets
template <class Ex>
int hs_lookahead_schedule(Ex& ex, int n, int P) {
    const int U0 = n + 1;
    int rc;
    // S2 must see everything already queued on S1 (e.g. the input assembly)
    if ((rc = ex.record(U0 + n, 1)) || (rc = ex.wait(2, U0 + n))) return rc;
    // POTRF and TRSM of panels k0 .. k0+P-1 on S2, right-looking inside the
    // step: panel k updates the step's later columns k+1 .. k0+P-1 at once
    auto chain = [&](int k0) -> int {
        int r;
        const int kend = std::min(n, k0 + P);
        for (int k = k0; k < kend; ++k) {
            if ((r = ex.begin(2, 0, k)) || (r = ex.potrf(k)) || (r = ex.end(2, 0))) return r;
            if (k + 1 < n && ((r = ex.begin(2, 1, k)) || (r = ex.trsm(k, k + 1, n - k - 1)) || (r = ex.end(2, 1))))
                return r;
            if (k + 1 < kend && ((r = ex.begin(2, 3, k)) ||
                                 (r = ex.trail(2, k, k + 2 == kend ? 5 : 6, k + 1, kend - k - 1, 1)) ||
                                 (r = ex.end(2, 3))))
                return r;
        }
        return HS_OK;
    };
    if ((rc = chain(0)) || (rc = ex.prep(0)) || (rc = ex.record(0, 2))) return rc;
    int step = 0;
    for (int k = 0; k + P < n; k += P, ++step) {
        if (step >= 1 && (rc = ex.wait(2, U0 + step - 1))) return rc;
        // columns k+P .. k+2P-1 by the step's P panels
        if ((rc = ex.begin(2, 3, k)) || (rc = ex.trail(2, k, P == 1 ? 5 : 6, k + P, P == 1 ? n - k - 1 : P, P)) ||
            (rc = ex.end(2, 3)))
            return rc;
        if ((rc = chain(k + P)) || (rc = ex.prep(k + P)) || (rc = ex.record(step + 1, 2))) return rc;
        if ((rc = ex.wait(1, step))) return rc;
        if (k + 2 * P < n) {
            if ((rc = ex.begin(1, 2, k)) || (rc = ex.trail(1, k, 4, k + 2 * P, n - k - 2 * P, P)) ||
                (rc = ex.end(1, 2)))
                return rc;
        }
        if ((rc = ex.record(U0 + step, 1))) return rc;
    }
    return ex.wait(1, step);
}

// The mixed factor's bf16 planes: the panels k0 .. k0+P-1 are split after
// their chain when a later step reads them, and the launches that read the
// split planes are the step's rest update and its P-panel column update.
inline bool hs_plane_prep_needed(int k0, int P, int n) { return k0 + P < n; }
inline bool hs_trail_reads_planes(int mode, int np, int P) {
    return np == P && (mode == 4 || (mode == 6 && P > 1));
}

struct NoPanelPrep {
    int operator()(hipStream_t, int) const { return HS_OK; }
};

// The executor that issues the schedule on the device: potrf(k) and
// trsm(k, blockStart, blockCount) launch on h->stream, which is S2 while they
// run; trail(stream, ...) and prep(stream, k0) take their stream.
template <class Potrf, class Trsm, class Trail, class Prep>
struct DeviceSchedule {
    hs_handle_t h;
    PhaseTimer& pt;
    Potrf& potrf_;
    Trsm& trsm_;
    Trail& trail_;
    Prep& prep_;
    hipStream_t s1, s2;
    hipStream_t str(int s) const { return s == 1 ? s1 : s2; }
    int record(int ev, int s) {
        HS_HIP(hipEventRecord(h->events[ev], str(s)));
        return HS_OK;
    }
    int wait(int s, int ev) {
        HS_HIP(hipStreamWaitEvent(str(s), h->events[ev], 0));
        return HS_OK;
    }
    int potrf(int k) {
        OnStream os(h, s2);
        return potrf_(k);
    }
    int trsm(int k, int start, int count) {
        OnStream os(h, s2);
        return trsm_(k, start, count);
    }
    int trail(int s, int k, int mode, int rowStart, int rowCount, int np) {
        return trail_(str(s), k, mode, rowStart, rowCount, np);
    }
    int prep(int k0) { return prep_(s2, k0); }
    int begin(int s, int phase, int k) { return pt.begin(str(s), phase, k); }
    int end(int s, int phase) { return pt.end(str(s), phase); }
};

// panels = 0 takes hs_schedule_panels(n); any other value is clamped to
// 1 .. n/2. prep(s2, k0) runs on S2 after the panels k0 .. k0+P-1 are
// factored and before the event that releases them to S1 (the mixed factor
// converts them to bf16 planes there, once for both streams).
template <class Potrf, class Trsm, class Trail, class Prep = NoPanelPrep>
int hs_lookahead_factor(hs_handle_t h, int n, PhaseTimer& pt, Potrf potrf, Trsm trsm, Trail trail, int panels = 0,
                        Prep prep = Prep()) {
    const int P = panels > 0 ? std::max(1, std::min(panels, n / 2)) : hs_schedule_panels(n);
    int rc = hs_ensure_events(h, 2 * (size_t)n + 2);
    if (rc) return rc;
    DeviceSchedule<Potrf, Trsm, Trail, Prep> ex{h, pt, potrf, trsm, trail, prep, h->stream, h->stream2};
    return hs_lookahead_schedule(ex, n, P);
}
