// nb_bh.hip -- Barnes-Hut pipeline on gfx950.
//
//   bbox      k_bbox_partial (per-block FP64 min/max), reduced to the root
//             cube by every block of the key kernel, exactly as
//             CPU/src/force_calculator_barnes_hut_optimized.cpp:105-124
//             (centre = (min+max)/2, size = max extent + 1e-9).
//   keys      k_keys: 63-bit octant-path keys, 21 levels of the reference's FP64
//             subdivision replayed per particle (child centre = c +/- (s/2)/2,
//             octant = x>=c | y>=c<<1 | z>=c<<2; same file :159-161,:219-227).
//             Sorting the keys orders particles by the CPU tree's child order.
//   sort      LSD radix sort of (key, index), 8 passes of 8 bits, stable: the
//             key kernel counts the digits of all passes, then one kernel per
//             pass (decoupled look-back across tiles, tile sorted in LDS, runs
//             written contiguously).
//   gather    sorted float4 (x, y, z, m), written by the last radix pass.
//   build     k_levels + k_col_scan + k_emit: the tree the reference's
//             insertion build produces (a cell splits iff it holds more than
//             leaf_capacity particles and its half size is >= 1e-9), built from
//             the sorted keys without host synchronisation.  A cell at level l
//             holding particle i has more than cap particles iff some window
//             [j, j+cap] with i-cap <= j <= i shares the level-l prefix, so each
//             particle's leaf level is 1 + max of the shared-prefix lengths of
//             key pairs (j, j+cap) around it.  Nodes are numbered in level
//             order (key order inside a level); children of a node are
//             contiguous.  Empty octants are not materialised.
//   up-sweep  k_leaf_moments (every leaf, from its particles), then k_upsweep
//             per level, deepest first (internal nodes, from their children):
//             mass, centre of mass and the traceless quadrupole in FP64 for
//             every level; FP32 copies for the walk.
//   walk      k_walk: persistent one-wave blocks take walk groups (the tree
//             cells of at most 128 particles) from an atomic counter, in
//             size classes largest first, key order inside a class (the order
//             is written by k_emit).  The wave
//             opens cells breadth-first (8 parents x 8 children per step, one
//             child per lane), accepting a cell for the whole group when
//             size^2 < theta^2 * d^2 with d the distance from the cell's centre
//             of mass to the group's bounding box (the reference criterion of
//             force_calculator_barnes_hut_optimized.cpp:88, evaluated
//             conservatively for every target of the group at once).  A node's
//             walk record is read once per candidate; accepted cells go from
//             registers to an LDS staging list, opened leaves' particles to a
//             second one, and each list is evaluated by all lanes (two targets
//             per lane, or three where that fills more lanes) in batches of
//             60-64 entries, a multiple of the number of lane subsets.  The first 64 frontier entries of a level are
//             kept in LDS.  An accepted cell far from the group compared with
//             the group's size (quadrupole walk) instead enters a third-order
//             local expansion about the group centre, evaluated once per
//             target.
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstddef>
#include <cstdlib>
#include <cstring>

#include "nb_internal.hpp"

namespace nb {

constexpr int kMaxDigits = 21;   // octant digits in a key
constexpr int kLevels = 22;      // node levels 0..21
constexpr int kBlk = 256;
constexpr int kDigitBits = 8;                  // radix digit: 8 bits, 8 passes cover the 63-bit keys
constexpr int kSortBins = 1 << kDigitBits;
constexpr int kSortPasses = (63 + kDigitBits - 1) / kDigitBits;
// Radix-pass tiles: a block of T threads holds R rows of 64 keys per wave,
// T * R keys.  Small systems take small tiles (more blocks to fill the chip),
// large ones large tiles (fewer look-back words per key).
constexpr int kKeyBlocks = 1024;     // grid cap of the key kernel (it strides)
constexpr int kKeyHistCopies = 4;    // copies of the digit table per key block
constexpr int kKeyBatch = 4;         // particles per thread per step of the key kernel
constexpr int kBuildRowsMax = 4;     // sorted particles per thread in the build kernels, at most
constexpr int kMomBatch = 8;  // particles per batch of loads in a leaf's moments
// Nodes of levels 0..kMomLevels-1 keep an FP64 moment record (mass, centre of
// mass, quadrupole), from which their parents' moments are accumulated; the
// nodes below keep only their FP32 walk records, and their parents read
// those.  Level order numbers the kept nodes first: at most kMomNodes.
constexpr int kMomLevels = 8;
constexpr int64_t kMomNodes = 2396745;  // sum of 8^l, l < kMomLevels

constexpr int kUpsweepBlocks = 2048;  // grid cap of a per-level up-sweep launch
constexpr int kBandThreads = 512;     // block of the single-block launch of a run of sparse levels
constexpr int kDenseLevel = 512;      // nodes that give a level its own up-sweep launch
constexpr int kBandCache = 8192;      // nodes whose child ranges that launch keeps in LDS
constexpr int kLookbackSpins = 1 << 22;  // polls before a tile gives up waiting (error bit 8)
// Walk frontier capacity per block (opened cells per level).  The frontier of
// a group grows like theta^-3: 2048 entries at theta = 0.5 (the largest seen on
// a 2^20 Plummer sphere is ~110), scaled for smaller theta up to kWalkFcapMax.
// Only the part a group uses is touched, so the reservation costs HBM
// capacity, not bandwidth.
constexpr int kWalkFcap = 2048;
constexpr int kWalkFcapMax = 65536;
constexpr int kWalkBlocksPerCU = 24;  // persistent one-wave walk blocks per CU (NB_WALK_BPC overrides)
constexpr int kWalkBlocksPerCUFar = 20;  // the same for the far-cell walk (5 waves per SIMD)
constexpr int kGroup = 128;     // largest walk group (a wave holds two or three targets per lane)
constexpr int kGroupDefault = 128;  // walk group: cells of at most 128 particles
constexpr int kFarBatch = 128;  // far cells per expansion batch: two per lane, one reduction
constexpr double kWalkFarRatio = 0.4;  // far cells: box half diagonal < 0.4 theta |R|
constexpr int64_t kMaxNodes = (int64_t)1 << 28;  // node pool limit (32-bit record offsets in the walk)
// Per-block count columns: 22 levels, group starts, leaf starts, oversize leaves.
// Walk groups are handed out by size class, largest first (a group's cost
// grows with its size: one interaction list evaluated for all its targets),
// in key order inside a class: with groups of at most g particles, classes
// >= 3g/4, g/2 .. 3g/4-1, g/4 .. g/2-1 and < g/4 particles.
constexpr int kGBuckets = 4;
constexpr int kColGroup = kLevels, kColLeaf = kLevels + 1, kColOver = kLevels + 2;
constexpr int kColGB = kLevels + 3;  // kGBuckets columns: group starts per size class
constexpr int kCols = kColGB + kGBuckets;

struct RootInfo {
    double center[3];
    double size;
    double minv[3], maxv[3];
    int lmax;         // deepest level a node can have (21-digit keys and the 1e-9 split rule)
    int key_limited;  // 1 if the split rule alone would allow deeper nodes than the keys encode
};

struct LevelInfo {
    int start[kLevels + 1];
    int count[kLevels];
    int total;
    int max_level;
    int n_groups;
    int gb_start[kGBuckets];  // first walk-order slot of each group size class
};

struct Counters {
    unsigned long long pp, pc;
    unsigned long long pf;   // far cells: (group, cell) pairs taken into a group's local expansion
    unsigned long long pft;  // (target, far cell) pairs those expansions stand for
    int max_frontier;
    int overflow;   // this evaluation: bit 0 walk frontier, bit 1 node capacity, bit 2
                    // malformed tree, bit 3 sort look-back wait limit, bit 4 non-finite input
    int oversize;   // leaves at the key-depth limit with > cap particles
    int n_leaves;
    int next_group;  // work counter of the persistent walk
    int sticky;      // OR of `overflow` over every evaluation since the flags were last read
                     // (not reset per evaluation; cleared by the host when it reads them)
};

// Radix sort state: the digit counts of all passes (filled in one pass over
// the keys) and each pass's tile counter.
struct SortState {
    int hist[kSortPasses][kSortBins];
    int tile[kSortPasses];
};

// Errors that make the sorted order (and so every later stage) invalid.
constexpr int kInputBad = 16;  // a non-finite position or mass
constexpr int kTreeBad = 2 | 4 | 8 | kInputBad;

}  // namespace nb

struct nb::BHWorkspace {
    int64_t cap_n = 0, cap_nodes = 0;
    int64_t last_n = 0;
    int last_cap = 8;
    int last_valid = 0;
    DevBuf root, bbox_part, keys0, keys1, vals0, vals1, hist, spos, upos;
    DevBuf lvl_leaf, lvl_sh, blkcnt, blkoff, levels, leaf_of, counters;
    DevBuf n_pstart, n_pend, n_child, n_cend, n_level, n_mom, w_mac, w_link, w_q0, w_q1;
    DevBuf walk_scratch, gflag, gstart, glev, totals, sortst, gorder, leaves;
    int walk_blocks = 0;  // blocks the frontier scratch is sized for
    int cus = 0;
    int group_size = 0;
    int sort_slots = 0, sort_slots_key = 0;  // radix-pass blocks resident at once, for which shape
    // Node counts per level of a recent evaluation, copied to pinned host
    // memory on a side stream while the walk runs; they plan the next
    // evaluation's up-sweep launches (a guide only: any plan is correct).
    hipStream_t hint_stream = nullptr;
    hipEvent_t hint_ready = nullptr, hint_done = nullptr;
    int* hint_host = nullptr;
    int hint[nb::kLevels] = {};
    bool have_hint = false, hint_pending = false;
    ~BHWorkspace() {
        if (hint_stream) {
            (void)hipStreamSynchronize(hint_stream);
            (void)hipStreamDestroy(hint_stream);
        }
        if (hint_ready) (void)hipEventDestroy(hint_ready);
        if (hint_done) (void)hipEventDestroy(hint_done);
        if (hint_host) (void)hipHostFree(hint_host);
    }
};

namespace nb {

// ------------------------------------------------------------------ helpers
__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ unsigned long long lanemask_lt() {
    return (1ull << lane_id()) - 1ull;
}

// v + (v of the lane the DPP control selects); lanes of rows outside
// kRowMask add 0.
template <int kCtrl, int kRowMask = 0xf>
__device__ __forceinline__ float dpp_add(float v) {
    return v + __int_as_float(
                   __builtin_amdgcn_update_dpp(0, __float_as_int(v), kCtrl, kRowMask, 0xf, false));
}

// keep + (send of the lane the DPP control selects): one step of a
// reduce-scatter in which paired lanes exchange the halves they give up.
template <int kCtrl>
__device__ __forceinline__ float dpp_pair_add(float keep, float send) {
    return keep + __int_as_float(
                      __builtin_amdgcn_update_dpp(0, __float_as_int(send), kCtrl, 0xf, 0xf, false));
}

__device__ __forceinline__ int wave_incl_scan(int v) {
    const int lane = lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    return v;
}

// Number of leading octant digits two 63-bit keys share (0..21).
__device__ __forceinline__ int shared_levels(unsigned long long a, unsigned long long b) {
    const unsigned long long d = a ^ b;
    if (d == 0) return kMaxDigits;
    return (__clzll((long long)d) - 1) / 3;
}

// ------------------------------------------------------------------ counters
__device__ __forceinline__ void reset_counters(Counters* c) {
    c->pp = 0;
    c->pc = 0;
    c->pf = 0;
    c->pft = 0;
    c->max_frontier = 0;
    c->overflow = 0;
    c->oversize = 0;
    c->n_leaves = 0;
    c->next_group = 0;
}

// Records error bits for this evaluation and in the sticky word.
__device__ __forceinline__ void raise_flag(Counters* c, int bits) {
    atomicOr(&c->overflow, bits);
    atomicOr(&c->sticky, bits);
}

// ------------------------------------------------------------------ bbox
// Per-block min/max of the positions; block 0 also clears this evaluation's
// counters and the radix sort's digit counts and tile counters (the key
// kernel that follows adds to them).
__global__ __launch_bounds__(kBlk) void k_bbox_partial(int n, const double* __restrict__ x,
                                                       const double* __restrict__ y,
                                                       const double* __restrict__ z,
                                                       double* __restrict__ part,
                                                       Counters* __restrict__ ctr,
                                                       SortState* __restrict__ ss) {
    __shared__ double sh[6][kBlk];
    if (blockIdx.x == 0) {
        if (threadIdx.x == 0) reset_counters(ctr);
#pragma unroll
        for (int e = threadIdx.x; e < kSortPasses * kSortBins; e += kBlk) (&ss->hist[0][0])[e] = 0;
        if (threadIdx.x < kSortPasses) ss->tile[threadIdx.x] = 0;
    }
    double mn[3] = {DBL_MAX, DBL_MAX, DBL_MAX};
    double mx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (int i = blockIdx.x * kBlk + threadIdx.x; i < n; i += gridDim.x * kBlk) {
        const double p[3] = {x[i], y[i], z[i]};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            mn[a] = p[a] < mn[a] ? p[a] : mn[a];
            mx[a] = p[a] > mx[a] ? p[a] : mx[a];
        }
    }
    const int t = threadIdx.x;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        sh[a][t] = mn[a];
        sh[3 + a][t] = mx[a];
    }
    __syncthreads();
    for (int s = kBlk / 2; s > 0; s >>= 1) {
        if (t < s) {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                sh[a][t] = fmin(sh[a][t], sh[a][t + s]);
                sh[3 + a][t] = fmax(sh[3 + a][t], sh[3 + a][t + s]);
            }
        }
        __syncthreads();
    }
    if (t < 6) part[blockIdx.x * 6 + t] = sh[t][0];
}

// Root cube from the nparts partial boxes, computed by every block of the
// key kernel (the min/max are exact, so every block gets the same root):
// centre = (min+max)/2, size = largest extent + 1e-9, and the deepest level
// a node can reach.  Needs sh[6][kBlk] and a block barrier per step.
__device__ __forceinline__ void reduce_root(int nparts, const double* __restrict__ part,
                                            double (*sh)[kBlk], RootInfo* r) {
    const int t = threadIdx.x;
    double v[6] = {DBL_MAX, DBL_MAX, DBL_MAX, -DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (int k = t; k < nparts; k += kBlk) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            v[a] = fmin(v[a], part[k * 6 + a]);
            v[3 + a] = fmax(v[3 + a], part[k * 6 + 3 + a]);
        }
    }
#pragma unroll
    for (int a = 0; a < 6; ++a) sh[a][t] = v[a];
    __syncthreads();
    for (int s = kBlk / 2; s > 0; s >>= 1) {
        if (t < s) {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                sh[a][t] = fmin(sh[a][t], sh[a][t + s]);
                sh[3 + a][t] = fmax(sh[3 + a][t], sh[3 + a][t + s]);
            }
        }
        __syncthreads();
    }
    if (t == 0) {
        double size = 0.0;
        for (int a = 0; a < 3; ++a) {
            r->minv[a] = sh[a][0];
            r->maxv[a] = sh[3 + a][0];
            r->center[a] = (sh[a][0] + sh[3 + a][0]) / 2;
            const double ext = sh[3 + a][0] - sh[a][0];
            size = a == 0 ? ext : (ext > size ? ext : size);
        }
        size += 1e-9;
        r->size = size;
        // A node of size s may split iff s/2 >= 1e-9
        // (force_calculator_barnes_hut_optimized.cpp:208-211).
        int l = 0;
        double s = size;
        while (l < kMaxDigits && s / 2.0 >= 1e-9) {
            s = s / 2.0;
            ++l;
        }
        r->lmax = l;
        r->key_limited = (l == kMaxDigits && s / 2.0 >= 1e-9) ? 1 : 0;
    }
    __syncthreads();
}

// ------------------------------------------------------------------ keys
// One pass over the FP64 state (grid-stride, at most kKeyBlocks blocks):
// every particle's 63-bit key and index, its FP32 (x, y, z, m) in original
// order (so the gather after the sort moves one 16-byte record per
// particle), and the digit counts of all eight radix passes.  The counts go
// into kKeyHistCopies copies of a block table in LDS (lane mod copies, rows
// padded by one word: the high digits of clustered inputs take few values)
// and are added to the global counts once per block.  Also zeroes the radix
// passes' look-back words.
__global__ __launch_bounds__(kBlk) void k_keys(int n, const double* __restrict__ x,
                                               const double* __restrict__ y,
                                               const double* __restrict__ z,
                                               const double* __restrict__ m, int nparts,
                                               const double* __restrict__ part,
                                               RootInfo* __restrict__ root,
                                               unsigned long long* __restrict__ keys,
                                               unsigned* __restrict__ vals,
                                               float4* __restrict__ upos,
                                               Counters* __restrict__ ctr,
                                               SortState* __restrict__ ss,
                                               unsigned* __restrict__ look, int nlook) {
    constexpr int kRow = kSortBins + 1;
    static_assert(kKeyHistCopies * kSortPasses * kRow * 4 >= 6 * kBlk * 8, "root scratch");
    __shared__ __align__(16) int h[kKeyHistCopies * kSortPasses * kRow];  // root scratch first
    __shared__ RootInfo sr;
    const int t = threadIdx.x;
    for (int i = blockIdx.x * kBlk + t; i < nlook; i += gridDim.x * kBlk) look[i] = 0u;
    reduce_root(nparts, part, reinterpret_cast<double(*)[kBlk]>(h), &sr);  // ends with a barrier
    if (blockIdx.x == 0 && t == 0) *root = sr;  // for the later stages
    for (int e = t; e < kKeyHistCopies * kSortPasses * kRow; e += kBlk) h[e] = 0;
    __syncthreads();
    int* hc = h + (t % kKeyHistCopies) * kSortPasses * kRow;
    const double rcx = sr.center[0], rcy = sr.center[1], rcz = sr.center[2];
    // Child-centre offsets per level: the subdivision halves the size twice
    // per level (half = s/2, offset = half/2), exact in binary, so the same
    // for every particle.
    double off[kMaxDigits];
    {
        double sz = sr.size;
#pragma unroll
        for (int l = 0; l < kMaxDigits; ++l) {
            const double half = sz / 2.0;
            off[l] = half / 2.0;
            sz = half;
        }
    }
    bool bad = false;
    // kKeyBatch particles per thread per step, all loads issued first.
    for (int i0 = blockIdx.x * kBlk * kKeyBatch; i0 < n; i0 += gridDim.x * kBlk * kKeyBatch) {
        double px[kKeyBatch], py[kKeyBatch], pz[kKeyBatch], pm[kKeyBatch];
#pragma unroll
        for (int b = 0; b < kKeyBatch; ++b) {
            const int i = min(i0 + b * kBlk + t, n - 1);
            px[b] = x[i];
            py[b] = y[i];
            pz[b] = z[i];
            pm[b] = m[i];
        }
#pragma unroll
        for (int b = 0; b < kKeyBatch; ++b) {
            const int i = i0 + b * kBlk + t;
            if (i < n) {
                upos[i] = make_float4((float)px[b], (float)py[b], (float)pz[b], (float)pm[b]);
                // A non-finite position or mass makes the evaluation invalid
                // (error bit 16): the up-sweep and walk skip it.
                bad = bad || !(isfinite(px[b]) && isfinite(py[b]) && isfinite(pz[b]) &&
                               isfinite(pm[b]));
                double cx = rcx, cy = rcy, cz = rcz;
                unsigned long long k = 0;
#pragma unroll
                for (int l = 0; l < kMaxDigits; ++l) {
                    unsigned d = 0;
                    if (px[b] >= cx) {
                        d |= 1u;
                        cx = cx + off[l];
                    } else {
                        cx = cx + (-off[l]);
                    }
                    if (py[b] >= cy) {
                        d |= 2u;
                        cy = cy + off[l];
                    } else {
                        cy = cy + (-off[l]);
                    }
                    if (pz[b] >= cz) {
                        d |= 4u;
                        cz = cz + off[l];
                    } else {
                        cz = cz + (-off[l]);
                    }
                    k = (k << 3) | d;
                }
                keys[i] = k;
                vals[i] = (unsigned)i;
#pragma unroll
                for (int p = 0; p < kSortPasses; ++p)
                    atomicAdd(&hc[p * kRow + (int)((k >> (kDigitBits * p)) & (kSortBins - 1))], 1);
            }
        }
    }
    const unsigned long long bb = __ballot(bad);  // one report per wave
    if (bb && lane_id() == __ffsll((long long)bb) - 1) raise_flag(ctr, kInputBad);
    __syncthreads();
    for (int e = t; e < kSortPasses * kSortBins; e += kBlk) {
        const int p = e / kSortBins, d = e % kSortBins;
        int c = 0;
#pragma unroll
        for (int q = 0; q < kKeyHistCopies; ++q) c += h[(q * kSortPasses + p) * kRow + d];
        if (c) atomicAdd(&ss->hist[p][d], c);
    }
}

// ------------------------------------------------------------------ radix sort
// LSD radix sort of (key, index), 8 passes of 8 bits, stable, one kernel per
// pass (single-pass "onesweep" scheme: digit counts of every pass are taken
// once up front; each tile finds the keys of the same digit in earlier tiles
// by decoupled look-back).
//
// Exclusive scan of every column of a column-major count table
// cnt[c * nrows + r] (one 1024-thread block per column, coalesced 1024-row
// tiles with a running carry); the column sums go to totals[c].
__global__ __launch_bounds__(1024) void k_col_scan(int nrows, int* __restrict__ cnt,
                                                   int* __restrict__ totals) {
    __shared__ int shw[16];
    const int t = threadIdx.x;
    const int w = t >> 6;
    int* col = cnt + (size_t)blockIdx.x * nrows;
    int carry = 0;
    // 4096 rows per step, four consecutive rows per thread.
    for (int r0 = 0; r0 < nrows; r0 += 4096) {
        const int r = r0 + 4 * t;
        int v[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) v[q] = r + q < nrows ? col[r + q] : 0;
        const int sum4 = v[0] + v[1] + v[2] + v[3];
        const int incl = wave_incl_scan(sum4);
        if (lane_id() == 63) shw[w] = incl;
        __syncthreads();
        int before = 0, tile = 0;
        for (int k = 0; k < 16; ++k) {
            const int sw = shw[k];
            before += k < w ? sw : 0;
            tile += sw;
        }
        int run = carry + before + incl - sum4;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (r + q < nrows) col[r + q] = run;
            run += v[q];
        }
        carry += tile;
        __syncthreads();
    }
    if (t == 0) totals[blockIdx.x] = carry;
}

// One pass: a block takes the next tile of kOsTile keys (tiles are handed
// out in launch order, so every tile a block waits for belongs to a block that
// is already running).  Wave w holds the w-th part of the tile in
// registers and ranks its keys per digit in (row, lane) order without block
// barriers; the per-wave counts give the tile's digit counts, which the
// block publishes (flag 1) before adding up the counts of earlier tiles,
// walking back until a tile with a published inclusive prefix (flag 2), and
// publishing its own inclusive prefix.  The tile is then sorted by digit in
// LDS (stable: wave, row, lane order is input order) and each digit's run is
// written contiguously.  Look-back word: flag << 30 | count (counts < 2^30).
template <int kOsThreads, int kSortRows, int kLookWindow>
__global__ __launch_bounds__(kOsThreads) void k_onesweep(
    int n, const unsigned long long* __restrict__ kin, const unsigned* __restrict__ vin,
    unsigned long long* __restrict__ kout, unsigned* __restrict__ vout, int pass,
    SortState* __restrict__ ss, unsigned* __restrict__ look, Counters* __restrict__ ctr,
    const float4* __restrict__ upos, float4* __restrict__ spos, int resident,
    unsigned* __restrict__ gtot) {
    constexpr int kOsTile = kOsThreads * kSortRows;
    constexpr int kWaves = kOsThreads / 64;
    static_assert(kOsThreads >= kSortBins, "one owner thread per digit");
    constexpr int kPerWave = kOsTile / kWaves;  // kSortRows rows of 64 keys
    __shared__ unsigned long long sk[kOsTile];  // keys sorted by digit, then values
    __shared__ int cw[kWaves][kSortBins];  // per-wave digit counts, then per-wave slot starts
    __shared__ int goff[kSortBins];        // output position of tile slot 0 of each digit
    __shared__ int sred[2][kSortBins / 64];
    __shared__ int stile, sbad;
    const int t = threadIdx.x;
    const int w = t >> 6;
    const int lane = lane_id();
    const int shift = kDigitBits * pass;
    // Tile index: with every tile of the grid resident at once (`resident`,
    // checked by the host against the occupancy) a tile only ever waits for
    // blocks that hold or will get a slot, so the block index serves;
    // otherwise tiles are handed out in start order by a ticket.
    // This pass's global count of digit t (owner threads), loaded up front.
    const int g = t < kSortBins ? ss->hist[pass][t] : 0;
    if (t == 0) {
        if (!resident) stile = atomicAdd(&ss->tile[pass], 1);
        sbad = 0;
    }
    for (int e = t; e < kWaves * kSortBins; e += kOsThreads) (&cw[0][0])[e] = 0;
    if (!resident) __syncthreads();
    const int tile = resident ? (int)blockIdx.x : stile;
    const int base = tile * kOsTile;
    const int tcnt = min(kOsTile, n - base);
    const int wbase = base + w * kPerWave + lane;

    unsigned long long key[kSortRows];
    unsigned val[kSortRows];
    int pos[kSortRows];
#pragma unroll
    for (int r = 0; r < kSortRows; ++r) {
        const int i = wbase + r * 64;
        key[r] = i < n ? kin[i] : 0ull;
        val[r] = i < n ? vin[i] : 0u;
    }
    __syncthreads();  // the zeroed counts (and sbad) are visible
    // Rank within the wave: lanes with the same digit found with one ballot per digit bit.
    const unsigned long long lt = lanemask_lt();
#pragma unroll
    for (int r = 0; r < kSortRows; ++r) {
        const bool valid = wbase + r * 64 < n;
        const int d = (int)((key[r] >> shift) & (kSortBins - 1));
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < kDigitBits; ++b) {
            const unsigned long long bb = __ballot((d >> b) & 1);
            peers &= ((d >> b) & 1) ? bb : ~bb;
        }
        const int rank = __popcll(peers & lt);
        const int before = cw[w][d];
        pos[r] = before + rank;
        __builtin_amdgcn_wave_barrier();
        if (valid && rank == 0) cw[w][d] = before + __popcll(peers);
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();

    // Thread t < 256 owns digit t: tile count, publication, offsets.
    const bool owner = t < kSortBins;
    int wc[kWaves];
    int c = 0, ic = 0, ig = 0;
    unsigned* my = look + (size_t)tile * kSortBins + (owner ? t : 0);
    if (owner) {
#pragma unroll
        for (int k = 0; k < kWaves; ++k) {
            wc[k] = cw[k][t];
            c += wc[k];
        }
        __hip_atomic_store(my, ((tile == 0 ? 2u : 1u) << 30) | (unsigned)c, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
        ic = wave_incl_scan(c);
        ig = wave_incl_scan(g);
        if (lane == 63) {
            sred[0][w] = ic;
            sred[1][w] = ig;
        }
    }
    __syncthreads();
    if (owner) {
        int loff = ic - c, dstart = ig - g;
        for (int k = 0; k < w; ++k) {
            loff += sred[0][k];
            dstart += sred[1][k];
        }
        int excl = 0;
        if (gtot) {
            // Every tile holds a slot: a two-level prefix.  The counts of this
            // tile's earlier group mates (groups of kLookWindow tiles), then the
            // totals of the earlier groups, each published by its group's last
            // tile once it has its own in-group prefix.  No tile waits on a
            // later one, so the waits end; each level takes a memory round
            // trip per kLookWindow words.
            const int grp = tile / kLookWindow, first = grp * kLookWindow;
            // Sum of cnt published words p0[j * kSortBins], j < cnt.
            auto sum_published = [&](const unsigned* p0, int cnt) {
                int sum = 0;
                for (int b0 = 0; b0 < cnt && !sbad; b0 += kLookWindow) {
                    const int m = min(kLookWindow, cnt - b0);
                    for (int spins = 0;; ++spins) {
                        unsigned v[kLookWindow];
#pragma unroll
                        for (int j = 0; j < kLookWindow; ++j)
                            v[j] = j < m ? __hip_atomic_load(p0 + (size_t)(b0 + j) * kSortBins,
                                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                         : (1u << 30);
                        bool ready = true;
                        int part = 0;
#pragma unroll
                        for (int j = 0; j < kLookWindow; ++j) {
                            ready = ready && (v[j] >> 30) != 0u;
                            part += (int)(v[j] & 0x3fffffffu);
                        }
                        if (ready) {
                            sum += part;
                            break;
                        }
                        if (spins > kLookbackSpins) {
                            raise_flag(ctr, 8);
                            sbad = 1;
                            break;
                        }
                        __builtin_amdgcn_s_sleep(1);
                    }
                }
                return sum;
            };
            const int in_grp = sum_published(look + (size_t)first * kSortBins + t, tile - first);
            if (tile == first + kLookWindow - 1 || tile == (int)gridDim.x - 1)
                __hip_atomic_store(gtot + (size_t)grp * kSortBins + t,
                                   (1u << 30) | (unsigned)(in_grp + c), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            excl = in_grp + sum_published(gtot + t, grp);
        } else if (tile > 0) {
            // Decoupled look-back for digit t, kLookWindow earlier tiles per
            // memory round trip: their words are consumed nearest first until a
            // tile with an inclusive prefix; an unpublished word stalls the walk,
            // which resumes at that tile.
            int tt = tile - 1;
            int spins = 0;
            bool done = false;
            while (!done) {
                unsigned v[kLookWindow];
#pragma unroll
                for (int j = 0; j < kLookWindow; ++j)
                    v[j] = __hip_atomic_load(look + (size_t)max(tt - j, 0) * kSortBins + t,
                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                int used = 0;
                bool stall = false;
#pragma unroll
                for (int j = 0; j < kLookWindow; ++j) {
                    if (!done && !stall) {
                        const unsigned f = v[j] >> 30;
                        if (f == 0u) {
                            stall = true;
                        } else {
                            excl += (int)(v[j] & 0x3fffffffu);
                            ++used;
                            done = f == 2u || tt - j == 0;
                        }
                    }
                }
                tt -= used;
                if (stall && !done) {
                    if (++spins > kLookbackSpins) {
                        raise_flag(ctr, 8);
                        sbad = 1;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
            __hip_atomic_store(my, (2u << 30) | (unsigned)(excl + c), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
        goff[t] = dstart + excl - loff;
        int st = loff;  // slot of wave k's first key of digit t
#pragma unroll
        for (int k = 0; k < kWaves; ++k) {
            cw[k][t] = st;
            st += wc[k];
        }
    }
    __syncthreads();

    // Sort the tile's keys by digit in LDS; each digit's keys of this tile are
    // one contiguous run in the output.  Then the same for the values, in the
    // same LDS (the destination of slot j is remembered from the key pass).
#pragma unroll
    for (int r = 0; r < kSortRows; ++r) {
        if (wbase + r * 64 < n) {
            pos[r] += cw[w][(int)((key[r] >> shift) & (kSortBins - 1))];
            sk[pos[r]] = key[r];
        }
    }
    __syncthreads();
    if (sbad) return;  // offsets unknown: leave the output alone (the error bit is set)
    int dst[kSortRows];
#pragma unroll
    for (int r = 0; r < kSortRows; ++r) {
        const int j = t + r * kOsThreads;
        dst[r] = -1;
        if (j < tcnt) {
            const unsigned long long k = sk[j];
            dst[r] = goff[(int)((k >> shift) & (kSortBins - 1))] + j;
            if ((unsigned)dst[r] < (unsigned)n) kout[dst[r]] = k;
        }
    }
    __syncthreads();
    unsigned* sv = reinterpret_cast<unsigned*>(sk);
#pragma unroll
    for (int r = 0; r < kSortRows; ++r)
        if (wbase + r * 64 < n) sv[pos[r]] = val[r];
    __syncthreads();
    if (spos) {
        // Last pass: the values are the sorted order, so the particles'
        // float4 records are gathered here too (spos[k] = upos[perm[k]]).
        unsigned vv[kSortRows];
        float4 pv[kSortRows];
#pragma unroll
        for (int r = 0; r < kSortRows; ++r) {
            vv[r] = (unsigned)dst[r] < (unsigned)n ? sv[t + r * kOsThreads] : 0u;
            pv[r] = upos[vv[r] < (unsigned)n ? vv[r] : 0u];
        }
#pragma unroll
        for (int r = 0; r < kSortRows; ++r)
            if ((unsigned)dst[r] < (unsigned)n) {
                vout[dst[r]] = vv[r];
                spos[dst[r]] = pv[r];
            }
        return;
    }
#pragma unroll
    for (int r = 0; r < kSortRows; ++r)
        if ((unsigned)dst[r] < (unsigned)n) vout[dst[r]] = sv[t + r * kOsThreads];
}

// ------------------------------------------------------------------ build
// Leaf level L_i and shared prefix with the predecessor sh_i of every sorted
// particle, the walk-group start flags, and per-block counts of nodes starting
// at each level (columns 0..21) and of group starts (column 22).  A block
// takes a tile of kRows rows of kBlk sorted particles, with one key halo and
// one set of block barriers for the whole tile (more rows: fewer blocks and
// halos for large n).
//
// A walk group is the tree cell of the smallest level that holds at most
// gsz (<= kGroup) particles (or the particle's leaf if that is shallower), so
// a group is spatially compact; its particles are contiguous in sorted order.
template <int kRows>
__global__ __launch_bounds__(kBlk) void k_levels(int n, const unsigned long long* __restrict__ keys,
                                                 int cap, int gsz, const RootInfo* __restrict__ root,
                                                 signed char* __restrict__ Lout,
                                                 signed char* __restrict__ shout,
                                                 unsigned char* __restrict__ gflag,
                                                 signed char* __restrict__ glev, int nblk,
                                                 int* __restrict__ blkcnt,
                                                 Counters* __restrict__ ctr) {
    constexpr int H = kGroup;       // key halo on both sides of the block's tile
    constexpr int T = kBlk * kRows;  // sorted particles per block, kRows rows of kBlk
    __shared__ unsigned long long sk[T + 2 * H];
    __shared__ signed char ss[T + H];
    __shared__ signed char sg[T + H];
    __shared__ int wcnt[4][kCols];  // per-wave counts, summed over the rows
    const int t = threadIdx.x;
    const int w = t >> 6;
    const int lane = lane_id();
    const int b0 = blockIdx.x * T;
    for (int q = t; q < T + 2 * H; q += kBlk) {
        const int g = b0 - H + q;
        sk[q] = (g >= 0 && g < n) ? keys[g] : 0ull;
    }
    if (lane < kCols) wcnt[w][lane] = 0;
    __syncthreads();
    // ss[u]: shared levels of keys (j, j+cap) with j = b0 - cap + u;
    // sg[u]: shared levels of keys (j, j+gsz) with j = b0 - gsz + u (gsz <= kGroup).
    for (int u = t; u < T + H; u += kBlk) {
        if (u < T + cap) {
            const int j = b0 - cap + u;
            ss[u] = (j >= 0 && j + cap < n)
                        ? (signed char)shared_levels(sk[H - cap + u], sk[H + u])
                        : (signed char)-1;
        }
        if (u < T + gsz) {
            const int jg = b0 - gsz + u;
            sg[u] = (jg >= 0 && jg + gsz < n)
                        ? (signed char)shared_levels(sk[H - gsz + u], sk[H + u])
                        : (signed char)-1;
        }
    }
    __syncthreads();
    // Window maxima of sg over gsz+1 entries by doubling in place: after step
    // k, sg[u] is the maximum of the original sg[u .. u + 2^k - 1] (entries
    // whose range would pass the end keep a partial value and are never read).
    constexpr int kPer = T / kBlk + 1;  // sg entries per thread (ng < kPer * kBlk)
    const int ng = T + gsz;
    int kg = 0;
    while ((2 << kg) <= gsz + 1) ++kg;
    for (int k = 0; k < kg; ++k) {
        const int h = 1 << k;
        int v[kPer];
#pragma unroll
        for (int m = 0; m < kPer; ++m) {
            const int u = t + m * kBlk;
            v[m] = u + 2 * h <= ng ? max((int)sg[u], (int)sg[u + h]) : 0;
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < kPer; ++m) {
            const int u = t + m * kBlk;
            if (u + 2 * h <= ng) sg[u] = (signed char)v[m];
        }
        __syncthreads();
    }
    const int lmax = root->lmax;
    const bool key_limited = root->key_limited != 0;
    for (int r = 0; r < kRows; ++r) {
        const int o = r * kBlk + t;  // position in the tile
        const int i = b0 + o;
        const bool valid = i < n;
        int Li = -1, shi = 0, mx = -1, gl = -1;
        if (valid) {
            for (int q = 0; q <= cap; ++q) mx = max(mx, (int)ss[o + q]);
            const int mg = max((int)sg[o], (int)sg[o + gsz + 1 - (1 << kg)]);
            Li = min(lmax, mx + 1);
            gl = min(Li, mg + 1);
            shi = (i == 0) ? -1 : shared_levels(sk[H + o - 1], sk[H + o]);
            Lout[i] = (signed char)Li;
            shout[i] = (signed char)shi;
            glev[i] = (signed char)gl;
        }
        const bool leaf_start = valid && shi < Li;
        const bool group_start = valid && shi < gl;
        int gb = -1;  // size class of the group starting here
        if (group_start) {
            // The group is the level-gl cell holding i: i and the particles
            // after it sharing its first gl octant digits, so it has more
            // than k particles iff particle i+k shares them.
            auto longer = [&](int k) {
                return i + k < n && shared_levels(sk[H + o], sk[H + o + k]) >= gl;
            };
            gb = longer(max(0, 3 * gsz / 4 - 1)) ? 0
                 : longer(max(0, gsz / 2 - 1)) ? 1
                 : longer(max(0, gsz / 4 - 1)) ? 2
                                                : 3;
        }
        if (valid) gflag[i] = (unsigned char)(gb + 1);  // 0: no group starts here
        const bool over = leaf_start && mx + 1 > lmax && key_limited;
        // Each wave's lane 0 adds its row's counts to the wave's own entries.
#pragma unroll
        for (int l = 0; l < kLevels; ++l) {
            const unsigned long long bal = __ballot(valid && shi < l && l <= Li);
            if (lane == 0) wcnt[w][l] += __popcll(bal);
        }
        const unsigned long long bg = __ballot(group_start), bl = __ballot(leaf_start),
                                 bo = __ballot(over);
        if (lane == 0) {
            wcnt[w][kColGroup] += __popcll(bg);
            wcnt[w][kColLeaf] += __popcll(bl);
            wcnt[w][kColOver] += __popcll(bo);
        }
#pragma unroll
        for (int b = 0; b < kGBuckets; ++b) {
            const unsigned long long bb = __ballot(gb == b);
            if (lane == 0) wcnt[w][kColGB + b] += __popcll(bb);
        }
    }
    __syncthreads();
    if (t < kCols)
        blkcnt[t * nblk + blockIdx.x] = wcnt[0][t] + wcnt[1][t] + wcnt[2][t] + wcnt[3][t];
}

// Level starts/counts, group and leaf totals from the column sums (one
// thread).  The node total is summed in 64 bits: 22 levels of up to n nodes
// each can pass 2^31.  Returns false for a tree larger than the pool (which
// is never above 2^28 nodes, the range of the walk's 32-bit record offsets);
// its total is stored clamped so every later stage sees total > node_cap and
// writes nothing.
__device__ bool level_info(const int* tot, int node_cap, LevelInfo* li) {
    long long run = 0;
    int ml = 0;
    for (int l = 0; l < kLevels; ++l) {
        li->start[l] = (int)min(run, (long long)INT_MAX);
        li->count[l] = tot[l];
        run += tot[l];
        if (tot[l] > 0) ml = l;
    }
    const int total = (int)min(run, (long long)INT_MAX);
    li->start[kLevels] = total;
    li->total = total;
    li->max_level = ml;
    li->n_groups = tot[kColGroup];
    int gs = 0;
    for (int b = 0; b < kGBuckets; ++b) {
        li->gb_start[b] = gs;
        gs += tot[kColGB + b];
    }
    return run <= (long long)node_cap;
}

// Writes the node records: particle range, level, child range; and the leaf
// of every sorted particle.  Particle i starts the nodes of levels shi+1..Li,
// ends those of levels shn+1..Li (shn: shared levels with i+1) and records
// its leaf at level Li, so a wave only visits the levels from the smallest of
// those bounds to its deepest leaf.  Every block derives the level layout
// from the column sums; block 0 publishes it (LevelInfo, leaf and group
// totals, error bit 2 for an overfull pool) for the later stages.
template <int kRows>
__global__ __launch_bounds__(kBlk) void k_emit(
    int n, const signed char* __restrict__ Lin, const signed char* __restrict__ shin, int nblk,
    const int* __restrict__ blkoff, const int* __restrict__ totals, LevelInfo* __restrict__ li,
    int node_cap, int* __restrict__ pstart, int* __restrict__ pend, int* __restrict__ child,
    int* __restrict__ cend, signed char* __restrict__ level, int* __restrict__ leaf_of,
    int* __restrict__ leaves, const unsigned char* __restrict__ gflag, int* __restrict__ gstart,
    int* __restrict__ gorder, Counters* __restrict__ ctr) {
    __shared__ int wcnt[4][kCols];  // per-wave counts of the current row
    __shared__ int rbase[kCols];    // per column: earlier blocks' and this block's earlier rows' counts
    __shared__ int stot[kCols];
    __shared__ LevelInfo sli;
    __shared__ int sfits;
    const int t = threadIdx.x;
    const int w = t >> 6;
    const int lane = lane_id();
    if (t < kCols) stot[t] = totals[t];
    __syncthreads();
    if (t == 0) {
        const bool fits = level_info(stot, node_cap, &sli);
        sfits = fits;
        if (blockIdx.x == 0) {
            *li = sli;
            gstart[sli.n_groups] = n;
            ctr->n_leaves = stot[kColLeaf];
            ctr->oversize = stot[kColOver];
            if (!fits) raise_flag(ctr, 2);
        }
    }
    if (t < kCols) rbase[t] = blkoff[t * nblk + blockIdx.x];
    __syncthreads();
    const int* lstart = sli.start;
    const unsigned long long lt = lanemask_lt();
    for (int r = 0; r < kRows; ++r) {
        const int i = blockIdx.x * (kBlk * kRows) + r * kBlk + t;
        const bool valid = i < n;
        const int Li = valid ? (int)Lin[i] : -1;
        const int shi = valid ? (int)shin[i] : kLevels;
        const int shn = (valid && i + 1 < n) ? (int)shin[i + 1] : -1;
        if (lane < kCols) wcnt[w][lane] = 0;
        // The wave's level range (lo, hi].
        int lo = valid ? min(min(shi, shn), Li - 1) : kLevels, hi = Li;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            lo = min(lo, __shfl_xor(lo, d, 64));
            hi = max(hi, __shfl_xor(hi, d, 64));
        }
        lo = max(lo, -1);
        __builtin_amdgcn_wave_barrier();
        for (int l = lo + 1; l <= hi; ++l) {
            const unsigned long long bal = __ballot(valid && shi < l && l <= Li);
            if (lane == 0) wcnt[w][l] = __popcll(bal);
        }
        const int gf = valid ? (int)gflag[i] : 0;
        const bool gs = gf != 0;
        const int gb = gf - 1;  // size class of the group starting here
        int grank, orank = 0;
        const bool first_of_leaf = valid && shi < Li;
        const unsigned long long bls = __ballot(first_of_leaf);
        if (lane == 0) wcnt[w][kColLeaf] = __popcll(bls);
        {
            const unsigned long long bal = __ballot(gs);
            grank = __popcll(bal & lt);
            if (lane == 0) wcnt[w][kColGroup] = __popcll(bal);
#pragma unroll
            for (int b = 0; b < kGBuckets; ++b) {
                const unsigned long long bb = __ballot(gb == b);
                if (gb == b) orank = __popcll(bb & lt);
                if (lane == 0) wcnt[w][kColGB + b] = __popcll(bb);
            }
        }
        __syncthreads();
        if (gs) {
            int base = rbase[kColGroup] + grank;
            for (int k = 0; k < w; ++k) base += wcnt[k][kColGroup];
            gstart[base] = i;
            // walk order: size classes largest first, key order inside a class
            int o = sli.gb_start[gb] + rbase[kColGB + gb] + orank;
            for (int k = 0; k < w; ++k) o += wcnt[k][kColGB + gb];
            gorder[o] = base;
        }
        if (sfits) {  // uniform: with the pool check failed, no node is written
            // Leaves in particle order (the leaf up-sweep reads their
            // particles in one sweep over the sorted array).
            int lrank = rbase[kColLeaf] + __popcll(bls & lt);
            for (int k = 0; k < w; ++k) lrank += wcnt[k][kColLeaf];
            // rank(l): nodes of level l started before this particle (all blocks).
            auto rank = [&](int l) {
                const unsigned long long bal = __ballot(valid && shi < l && l <= Li);
                int rk = rbase[l] + __popcll(bal & lt);
                for (int k = 0; k < w; ++k) rk += wcnt[k][l];
                return rk;
            };
            int rl = lo + 1 >= 0 && lo + 1 < kLevels ? rank(lo + 1) : 0;
            for (int l = max(lo + 1, 0); l <= hi; ++l) {
                const int rn = l + 1 < kLevels ? rank(l + 1) : 0;  // the same at level l+1
                if (valid && l <= Li && (l == Li || l > min(shi, shn))) {
                    const bool starts = shi < l;
                    const int id = lstart[l] + rl + (starts ? 1 : 0) - 1;
                    if (id < lstart[l] || id >= lstart[l + 1]) {
                        raise_flag(ctr, 4);
                    } else {
                        if (starts) {
                            pstart[id] = i;
                            level[id] = (signed char)l;
                            child[id] = (l < Li) ? lstart[l + 1] + rn : -1;
                        }
                        if (l > shn) {
                            pend[id] = i + 1;
                            cend[id] = (l < Li) ? lstart[l + 1] + rn + (shi <= l ? 1 : 0) : -1;
                        }
                        if (l == Li) {
                            leaf_of[i] = id;
                            if (first_of_leaf) leaves[lrank] = id;
                        }
                    }
                }
                rl = rn;
            }
        }
        __syncthreads();  // every read of this row's counts is done
        if (t < kCols) rbase[t] += wcnt[0][t] + wcnt[1][t] + wcnt[2][t] + wcnt[3][t];
        __syncthreads();
    }
}


// ------------------------------------------------------------------ up-sweep
// Moments of node `id` at level `lev` (FP64): a leaf (kLeaf) from its
// particles [c0, c1), an internal node from its (already finished) children
// [c0, c1) in child order.  Writes the FP64 record and the FP32 walk copies.
template <bool kLeaf>
__device__ __forceinline__ void node_moments(int id, int c0, int c1, int lev, int n, int total,
                                             const RootInfo* root, const float4* __restrict__ spos,
                                             double* mom, float4* mac, int2* link, float4* q0,
                                             float4* q1, double inv_theta2, int quad,
                                             Counters* ctr) {
    // Structural checks: a malformed tree is reported, never dereferenced.
    if constexpr (kLeaf) {
        if (c0 < 0 || c1 > n || c0 >= c1) {
            raise_flag(ctr, 4);
            link[id] = make_int2(~0, 0);
            return;
        }
    } else {
        if (c0 <= id || c1 > total || c0 >= c1 || c1 - c0 > 8 || lev < 0 || lev >= kLevels) {
            raise_flag(ctr, 4);
            link[id] = make_int2(~0, 0);
            return;
        }
    }
    // Every load of a batch is issued before the first is used: indices past
    // the end repeat the last particle (child) and are not counted.  The sums
    // run in particle (child) order.
    double M = 0, cx = 0, cy = 0, cz = 0;
    double Q[6] = {0, 0, 0, 0, 0, 0};  // xx yy zz xy xz yz
    if constexpr (kLeaf) {
        const int a = c0, b = c1;
        double sx = 0, sy = 0, sz = 0, ux = 0, uy = 0, uz = 0;
        for (int q0 = a; q0 < b; q0 += kMomBatch) {
            float4 pb[kMomBatch];
#pragma unroll
            for (int k = 0; k < kMomBatch; ++k) pb[k] = spos[min(q0 + k, b - 1)];
#pragma unroll
            for (int k = 0; k < kMomBatch; ++k) {
                if (q0 + k < b) {
                    const double m = pb[k].w;
                    M += m;
                    sx += m * pb[k].x;
                    sy += m * pb[k].y;
                    sz += m * pb[k].z;
                    ux += pb[k].x;
                    uy += pb[k].y;
                    uz += pb[k].z;
                }
            }
        }
        if (M > 0) {
            cx = sx / M;
            cy = sy / M;
            cz = sz / M;
        } else {
            const double inv = 1.0 / (double)(b - a);
            cx = ux * inv;
            cy = uy * inv;
            cz = uz * inv;
        }
        if (quad) {
            for (int q0 = a; q0 < b; q0 += kMomBatch) {
                float4 pb[kMomBatch];
#pragma unroll
                for (int k = 0; k < kMomBatch; ++k) pb[k] = spos[min(q0 + k, b - 1)];
#pragma unroll
                for (int k = 0; k < kMomBatch; ++k) {
                    if (q0 + k < b) {
                        const double m = pb[k].w;
                        const double dx = pb[k].x - cx, dy = pb[k].y - cy, dz = pb[k].z - cz;
                        const double r2 = dx * dx + dy * dy + dz * dz;
                        Q[0] += m * (3.0 * dx * dx - r2);
                        Q[1] += m * (3.0 * dy * dy - r2);
                        Q[2] += m * (3.0 * dz * dz - r2);
                        Q[3] += m * (3.0 * dx * dy);
                        Q[4] += m * (3.0 * dx * dz);
                        Q[5] += m * (3.0 * dy * dz);
                    }
                }
            }
        }
    } else {
        // At most 8 children (checked above): one batch.  Children of the
        // top levels have FP64 records (M, cx | cy, cz | Qxx, Qyy | Qzz, Qxy |
        // Qxz, Qyz), 16-byte aligned; deeper children only their FP32 walk
        // records, read into the same form.
        double2 rc[8][5];
        if (lev + 1 < kMomLevels) {
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const double2* r =
                    reinterpret_cast<const double2*>(mom + (size_t)min(c0 + k, c1 - 1) * 10);
#pragma unroll
                for (int j = 0; j < 5; ++j) rc[k][j] = r[j];
            }
        } else {
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const int c = min(c0 + k, c1 - 1);
                const float4 m4 = mac[c], a4 = q0[c], b4 = q1[c];
                rc[k][0] = make_double2(a4.x, m4.x);
                rc[k][1] = make_double2(m4.y, m4.z);
                rc[k][2] = make_double2(a4.y, a4.z);
                rc[k][3] = make_double2(a4.w, b4.x);
                rc[k][4] = make_double2(b4.y, b4.z);
            }
        }
        double sx = 0, sy = 0, sz = 0, ux = 0, uy = 0, uz = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if (c0 + k < c1) {
                const double m = rc[k][0].x;
                M += m;
                sx += m * rc[k][0].y;
                sy += m * rc[k][1].x;
                sz += m * rc[k][1].y;
                ux += rc[k][0].y;
                uy += rc[k][1].x;
                uz += rc[k][1].y;
            }
        }
        if (M > 0) {
            cx = sx / M;
            cy = sy / M;
            cz = sz / M;
        } else {
            const double inv = 1.0 / (double)(c1 - c0);
            cx = ux * inv;
            cy = uy * inv;
            cz = uz * inv;
        }
        if (quad) {
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                if (c0 + k < c1) {
                    const double m = rc[k][0].x;
                    const double dx = rc[k][0].y - cx, dy = rc[k][1].x - cy, dz = rc[k][1].y - cz;
                    const double r2 = dx * dx + dy * dy + dz * dz;
                    Q[0] += rc[k][2].x + m * (3.0 * dx * dx - r2);
                    Q[1] += rc[k][2].y + m * (3.0 * dy * dy - r2);
                    Q[2] += rc[k][3].x + m * (3.0 * dz * dz - r2);
                    Q[3] += rc[k][3].y + m * (3.0 * dx * dy);
                    Q[4] += rc[k][4].x + m * (3.0 * dx * dz);
                    Q[5] += rc[k][4].y + m * (3.0 * dy * dz);
                }
            }
        }
    }
    if (lev < kMomLevels) {  // id < kMomNodes: level order numbers these first
        double2* mo = reinterpret_cast<double2*>(mom + (size_t)id * 10);
        mo[0] = make_double2(M, cx);
        mo[1] = make_double2(cy, cz);
        mo[2] = make_double2(Q[0], Q[1]);
        mo[3] = make_double2(Q[2], Q[3]);
        mo[4] = make_double2(Q[4], Q[5]);
    }
    const double size = ldexp(root->size, -lev);
    const double r2open = size * size * inv_theta2;
    mac[id] = make_float4((float)cx, (float)cy, (float)cz, (float)r2open);
    link[id] = kLeaf ? make_int2(~c0, c1) : make_int2(c0, c1);
    q0[id] = make_float4((float)M, (float)Q[0], (float)Q[1], (float)Q[2]);
    q1[id] = make_float4((float)Q[3], (float)Q[4], (float)Q[5], 0.f);
}

// Leaves depend on their particles only: all of them in one launch, in
// particle order (the leaf list of k_emit), so neighbouring threads read
// neighbouring particle ranges.
__global__ __launch_bounds__(kBlk) void k_leaf_moments(
    const LevelInfo* __restrict__ li, const RootInfo* __restrict__ root,
    const int* __restrict__ pstart, const int* __restrict__ pend, const int* child,
    const int* cend, const signed char* __restrict__ level, const float4* __restrict__ spos,
    double* mom, float4* mac, int2* link, float4* q0, float4* q1, double inv_theta2, int quad,
    int n, const int* __restrict__ leaves, Counters* ctr) {
    if (ctr->overflow & kTreeBad) return;
    const int total = li->total;
    const int nl = ctr->n_leaves;
    for (int k = blockIdx.x * kBlk + threadIdx.x; k < nl; k += gridDim.x * kBlk) {
        const int id = leaves[k];
        if ((unsigned)id >= (unsigned)total || child[id] >= 0) {
            raise_flag(ctr, 4);
            continue;
        }
        node_moments<true>(id, pstart[id], pend[id], level[id], n, total, root, spos, mom, mac,
                           link, q0, q1, inv_theta2, quad, ctr);
    }
}


// Internal nodes, level-synchronous: one launch per level, deepest level
// first, so every child is finished (and visible at the kernel boundary)
// before its parent is computed.  (A single-launch version with per-node
// arrival counters needs a device-scope release/acquire per node, which on
// this GPU writes back / invalidates the XCD-local L2 and was 5x slower.)
__global__ __launch_bounds__(kBlk) void k_upsweep(
    int l, const LevelInfo* __restrict__ li, const RootInfo* __restrict__ root,
    const int* __restrict__ pstart, const int* __restrict__ pend, const int* child,
    const int* cend, const signed char* __restrict__ level, const float4* __restrict__ spos,
    double* mom, float4* mac, int2* link, float4* q0, float4* q1, double inv_theta2, int quad,
    int n, Counters* ctr) {
    if (ctr->overflow & kTreeBad) return;
    const int cnt = li->count[l];
    const int st = li->start[l];
    const int total = li->total;
    for (int k = blockIdx.x * kBlk + threadIdx.x; k < cnt; k += gridDim.x * kBlk) {
        const int c0 = child[st + k];
        if (c0 >= 0)
            node_moments<false>(st + k, c0, cend[st + k], l, n, total, root, spos, mom, mac, link,
                                q0, q1, inv_theta2, quad, ctr);
    }
}

// A run of consecutive sparse levels, lhi down to llo, in one block: the
// block barrier between levels replaces a kernel boundary (the block's own
// stores are visible to it after the barrier).
__global__ __launch_bounds__(kBandThreads) void k_upsweep_band(
    int lhi, int llo, const LevelInfo* __restrict__ li, const RootInfo* __restrict__ root,
    const int* __restrict__ pstart, const int* __restrict__ pend, const int* child,
    const int* cend, const signed char* __restrict__ level, const float4* __restrict__ spos,
    double* mom, float4* mac, int2* link, float4* q0, float4* q1, double inv_theta2, int quad,
    int n, Counters* ctr) {
    // The run's nodes are one id range (level order); their child ranges are
    // read into LDS once, up front, when the range fits, so a level costs one
    // round trip to memory (the children's records) instead of three.
    __shared__ int sc0[kBandCache], sc1[kBandCache];
    __shared__ int bad;
    const int t = threadIdx.x;
    if (t == 0) bad = ctr->overflow & kTreeBad;
    const int total = li->total;
    const int r0 = li->start[llo], r1 = li->start[lhi + 1];
    const bool cached = r1 - r0 <= kBandCache;
    if (cached) {
        constexpr int kPer = kBandCache / kBandThreads;
        int v0[kPer], v1[kPer];
#pragma unroll
        for (int j = 0; j < kPer; ++j) {  // every load in flight before the first store
            const int k = t + j * kBandThreads;
            v0[j] = k < r1 - r0 ? child[r0 + k] : 0;
            v1[j] = k < r1 - r0 ? cend[r0 + k] : 0;
        }
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const int k = t + j * kBandThreads;
            if (k < r1 - r0) {
                sc0[k] = v0[j];
                sc1[k] = v1[j];
            }
        }
    }
    __syncthreads();
    if (bad) return;
    for (int l = lhi; l >= llo; --l) {
        const int cnt = li->count[l];
        const int st = li->start[l];
        for (int k = t; k < cnt; k += kBandThreads) {
            const int id = st + k;
            const int c0 = cached ? sc0[id - r0] : child[id];
            if (c0 >= 0)
                node_moments<false>(id, c0, cached ? sc1[id - r0] : cend[id], l, n, total, root,
                                    spos, mom, mac, link, q0, q1, inv_theta2, quad, ctr);
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------ walk
// Persistent: each one-wave block takes groups of 64 sorted targets from an
// atomic counter.  The breadth-first frontier holds the child ranges of
// opened cells and lives in a per-block slice of global scratch (it can reach
// thousands of entries in dense regions); interaction lists live in LDS.
//
// Far cells (kFar): an accepted cell whose centre of mass is far from the
// group's box centre c compared with the box's half diagonal h
// (h^2 < rho2 |R|^2, R = com - c) is not evaluated at every target.  Its field
// enters a local Taylor expansion of the acceleration about c, to third order:
//   a(c + d) = L1 - L2.d + 1/2 L3:dd - 1/6 L4:ddd,
// with L1 the softened monopole + quadrupole acceleration at c, L2 the
// monopole + quadrupole field gradient and L3, L4 the monopole's higher
// derivatives (-M d^n(1/r)/dR^n, traceless, unsoftened).  L2..L4 are stored
// scaled by h^(n-1) against d/h, so every stored number is bounded by the
// cell's M/r^2.  Far cells are staged by node id, 128 at a time; each lane
// adds up two cells' 24 coefficients (3 + 5 + 7 + 9 independent components),
// a fixed-order reduce-scatter sums them over the wave and eight lanes add the
// sums to the group's LDS accumulators.  Each target evaluates the expansion
// once at the end.  The far-cell walk runs 5 waves per SIMD (96 VGPRs): the
// extra registers pay for the third target and the second far cell per lane.
template <bool kQuad, bool kEps0, bool kFar>
__global__ __launch_bounds__(64, kFar ? 5 : 6) void k_walk(int n, const float4* __restrict__ spos,
                                             const unsigned* __restrict__ perm,
                                             const float4* __restrict__ mac,
                                             const int2* __restrict__ link,
                                             const float4* __restrict__ q0,
                                             const float4* __restrict__ q1, float soft2,
                                             double G, double* __restrict__ ax,
                                             double* __restrict__ ay, double* __restrict__ az,
                                             int2* __restrict__ scratch, int fcap,
                                             const int* __restrict__ gstart,
                                             const int* __restrict__ gorder,
                                             const LevelInfo* __restrict__ li,
                                             Counters* __restrict__ ctr, float rho2) {
    // A tree that failed the input, sort, build or up-sweep checks is not
    // walked; this evaluation's accelerations are NaN rather than stale.
    if (ctr->overflow & kTreeBad) {
        const double nan = __builtin_nan("");
        for (int i = blockIdx.x * 64 + threadIdx.x; i < n; i += gridDim.x * 64) {
            ax[i] = nan;
            ay[i] = nan;
            az[i] = nan;
        }
        return;
    }
    __shared__ float4 pl[128];                 // staged source particles
    __shared__ float4 sc0[64], sc1[64], sc2[64];  // staged cells: centre of mass, (M Qxx Qyy Qzz), (Qxy Qxz Qyz)
    constexpr int kFl = 64;
    __shared__ int2 fl[2][kFl];                // frontier entries 0..63 of this level / the next
    __shared__ int flist[kFar ? kFarBatch : 1];  // staged far cells (node ids)
    __shared__ float lexp[kFar ? 24 : 1];      // the group's local expansion coefficients
    __shared__ int sgroup;

    const int lane = threadIdx.x;
    const unsigned long long lt = lanemask_lt();
    // Frontier entries from kFl on live in a per-block slice of global scratch.
    int2* fr0 = scratch + (size_t)blockIdx.x * 2 * fcap;
    int2* fr1 = fr0 + fcap;
    const int ngroups = li->n_groups;
    unsigned long long tot_pp = 0, tot_pc = 0, tot_pf = 0, tot_pft = 0;
    int maxf = 0;
    bool ovf = false;

    for (;;) {
        if (lane == 0) {
            const int k = atomicAdd(&ctr->next_group, 1);
            sgroup = k < ngroups ? gorder[k] : ngroups;
        }
        __syncthreads();
        const int grp = sgroup;
        __syncthreads();
        if (grp >= ngroups) break;
        const int gend = gstart[grp + 1];
      for (int g0 = gstart[grp]; g0 < gend; g0 += 128) {  // oversize leaves in chunks of 128
        const int nact = min(128, gend - g0);
        // Every lane holds two targets, A = tix and B = tix + nl, of a subset of
        // nl = ceil(nact/2) lanes, or (far-cell walk, t3) three, + C = tix + 2 nl,
        // of ceil(nact/3) lanes where that keeps more (target, entry) slots busy:
        // nsub3 nact of 192 against nsub2 nact of 128 per trip.  The wave runs
        // nsub = 64/nl such subsets; each takes every nsub-th entry of the
        // interaction lists and the partial sums are added in subset order at
        // the end.
        const int nl2 = (nact + 1) >> 1, nl3 = (nact + 2) / 3;
        const bool t3 = kFar && 2 * (64 / nl3) > 3 * (64 / nl2);
        const int nl = t3 ? nl3 : nl2;
        const int nsub = 64 / nl;
        const int cfull = nsub * (64 / nsub);  // staged cells (particles) per evaluation batch, >= 60
        const int sub = lane / nl;
        const int tix = lane - sub * nl;
        const bool live = sub < nsub;
        const bool hasB = live && tix + nl < nact;
        const bool hasC = t3 && live && tix + 2 * nl < nact;
        const bool lead = live && sub == 0;
        const int kA = g0 + tix, kB = g0 + tix + nl, kC = g0 + tix + 2 * nl;
        const float4 meA = spos[live ? kA : g0];
        const float4 meB = hasB ? spos[kB] : meA;  // a lane without B computes A twice
        const float4 meC = hasC ? spos[kC] : meA;  // likewise C

        // Group bounding box.
        float lo0 = fminf(fminf(meA.x, meB.x), meC.x), lo1 = fminf(fminf(meA.y, meB.y), meC.y),
              lo2 = fminf(fminf(meA.z, meB.z), meC.z);
        float hi0 = fmaxf(fmaxf(meA.x, meB.x), meC.x), hi1 = fmaxf(fmaxf(meA.y, meB.y), meC.y),
              hi2 = fmaxf(fmaxf(meA.z, meB.z), meC.z);
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            lo0 = fminf(lo0, __shfl_xor(lo0, d, 64));
            lo1 = fminf(lo1, __shfl_xor(lo1, d, 64));
            lo2 = fminf(lo2, __shfl_xor(lo2, d, 64));
            hi0 = fmaxf(hi0, __shfl_xor(hi0, d, 64));
            hi1 = fmaxf(hi1, __shfl_xor(hi1, d, 64));
            hi2 = fmaxf(hi2, __shfl_xor(hi2, d, 64));
        }
        const float bc0 = 0.5f * (lo0 + hi0), bc1 = 0.5f * (lo1 + hi1), bc2 = 0.5f * (lo2 + hi2);
        const float bh0 = fmaxf(hi0 - bc0, bc0 - lo0), bh1 = fmaxf(hi1 - bc1, bc1 - lo1),
                    bh2 = fmaxf(hi2 - bc2, bc2 - lo2);
        // Far-cell test: h^2 < rho2 |R|^2 with h the box's half diagonal.
        const float rg2 = fmaf(bh0, bh0, fmaf(bh1, bh1, bh2 * bh2));
        // (A box below 1e-30 counts as a point: its orders 1-3 vanish against order 0.)
        const float hg = sqrtf(rg2), hinv = hg > 1e-30f ? 1.f / hg : 0.f;
        int nfar = 0;
        unsigned long long npf = 0;
        if (kFar) {
            __syncthreads();  // the previous group's expansion has been read
            if (lane < 24) lexp[lane] = 0.f;
        }

        float fxA = 0.f, fyA = 0.f, fzA = 0.f, fxB = 0.f, fyB = 0.f, fzB = 0.f;
        float fxC = 0.f, fyC = 0.f, fzC = 0.f;
        int ncl = 0, npl = 0, nnext = 0, nb = 0;
        bool gov = false;  // this group's frontier overflowed: its sums are incomplete
        unsigned long long npp = 0, npc = 0;
        int2* nxtg = fr0;

        auto pp_one = [&](const float4& s, const float4& me, float& fx, float& fy, float& fz) {
            const float dx = s.x - me.x, dy = s.y - me.y, dz = s.z - me.z;
            const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, soft2)));
            float inv = __builtin_amdgcn_rsqf(r2);
            if (kEps0) inv = r2 > 0.f ? inv : 0.f;
            const float c = (s.w * inv) * (inv * inv);
            fx = fmaf(dx, c, fx);
            fy = fmaf(dy, c, fy);
            fz = fmaf(dz, c, fz);
        };

        // Subset sub evaluates slots sub, sub + nsub, ... of a batch.  Every
        // subset makes the same number of trips, so the loop counter stays
        // scalar: a partial last batch is padded to a multiple of nsub with
        // massless copies of slot 0, which add exact zeros.  Lanes outside the
        // subsets follow subset 0 and their sums are never read.
        const int jfirst = live ? sub : 0;
        auto trips_for = [&](int cnt) {
            return __builtin_amdgcn_readfirstlane(cnt == cfull ? cfull / nsub
                                                                : (cnt + nsub - 1) / nsub);
        };

        auto flush_particles = [&](int cnt) {
            const int trips = trips_for(cnt);
            if (trips * nsub > cnt) {  // only a group's last batch: cnt < cfull <= 64
                float4 v = pl[0];
                v.w = 0.f;
                if (lane >= cnt && lane < trips * nsub) pl[lane] = v;
                __syncthreads();
            }
            int j = jfirst;
            if (t3) {
                for (int t = 0; t < trips; ++t, j += nsub) {
                    const float4 s = pl[j];
                    pp_one(s, meA, fxA, fyA, fzA);
                    pp_one(s, meB, fxB, fyB, fzB);
                    pp_one(s, meC, fxC, fyC, fzC);
                }
            } else {
                for (int t = 0; t < trips; ++t, j += nsub) {
                    const float4 s = pl[j];
                    pp_one(s, meA, fxA, fyA, fzA);
                    pp_one(s, meB, fxB, fyB, fzB);
                }
            }
            npp += (unsigned long long)cnt;
        };

        // Quadrupole term: Q = [[xx xy xz],[xy yy yz],[xz yz zz]] with
        // qa = (M, xx, yy, zz), qb = (xy, xz, yz).
        auto pc_one = [&](const float4& c, const float4& qa, const float4& qb, const float4& me,
                          float& fx, float& fy, float& fz) {
            const float dx = c.x - me.x, dy = c.y - me.y, dz = c.z - me.z;
            const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, soft2)));
            const float inv = __builtin_amdgcn_rsqf(r2);
            const float inv2 = inv * inv;
            if (kQuad) {
                // a = d (M/r^3 + 5/2 dQd/r^7) - Qd/r^5
                //   = (1/r^5) [d (M r^2 + 5/2 dQd/r^2) - Qd]
                const float qdx = fmaf(qa.y, dx, fmaf(qb.x, dy, qb.y * dz));
                const float qdy = fmaf(qb.x, dx, fmaf(qa.z, dy, qb.z * dz));
                const float qdz = fmaf(qb.y, dx, fmaf(qb.z, dy, qa.w * dz));
                const float dqd = fmaf(dx, qdx, fmaf(dy, qdy, dz * qdz));
                const float inv5 = (inv2 * inv2) * inv;
                const float sc = fmaf(2.5f, dqd * inv2, qa.x * r2);
                fx = fmaf(fmaf(dx, sc, -qdx), inv5, fx);
                fy = fmaf(fmaf(dy, sc, -qdy), inv5, fy);
                fz = fmaf(fmaf(dz, sc, -qdz), inv5, fz);
            } else {
                const float cm = qa.x * (inv * inv2);
                fx = fmaf(dx, cm, fx);
                fy = fmaf(dy, cm, fy);
                fz = fmaf(dz, cm, fz);
            }
        };

        // Evaluates the cnt staged cells; the staging slots are free afterwards.
        auto flush_cells = [&](int cnt) {
            const int trips = trips_for(cnt);
            if (trips * nsub > cnt) {  // only a group's last batch: cnt < cfull <= 64
                __syncthreads();
                const float4 v = sc0[0];  // an accepted cell: r^2 > 0 for every target
                const float4 zero = make_float4(0.f, 0.f, 0.f, 0.f);
                if (lane >= cnt && lane < trips * nsub) {
                    sc0[lane] = v;
                    sc1[lane] = zero;
                    if (kQuad) sc2[lane] = zero;
                }
            }
            __syncthreads();
            int j = jfirst;
            if (t3) {
                for (int t = 0; t < trips; ++t, j += nsub) {
                    const float4 c = sc0[j];
                    const float4 qa = sc1[j];
                    const float4 qb = kQuad ? sc2[j] : qa;
                    pc_one(c, qa, qb, meA, fxA, fyA, fzA);
                    pc_one(c, qa, qb, meB, fxB, fyB, fzB);
                    pc_one(c, qa, qb, meC, fxC, fyC, fzC);
                }
            } else {
                for (int t = 0; t < trips; ++t, j += nsub) {
                    const float4 c = sc0[j];
                    const float4 qa = sc1[j];
                    const float4 qb = kQuad ? sc2[j] : qa;
                    pc_one(c, qa, qb, meA, fxA, fyA, fzA);
                    pc_one(c, qa, qb, meB, fxB, fyB, fzB);
                }
            }
            npc += (unsigned long long)cnt;
            __syncthreads();
        };

        // Adds the local expansion terms of the cnt (<= kFarBatch) staged far
        // cells to lexp: lane l sums cells l and l + 64 (a lane past cnt
        // evaluates a massless cell at unit distance, which adds exact zeros),
        // then the wave sums the lanes.  The staging slots are free afterwards.
        auto flush_far = [&](int cnt) {
            __syncthreads();
            const int trips = __builtin_amdgcn_readfirstlane((cnt + 63) >> 6);
            int fids[kFarBatch / 64];
#pragma unroll
            for (int u = 0; u < kFarBatch / 64; ++u) {
                const int slot = u * 64 + lane;
                fids[u] = slot < cnt ? flist[slot] : -1;
            }
            __syncthreads();
            float c[24];
#pragma unroll
            for (int j = 0; j < 24; ++j) c[j] = 0.f;
#pragma unroll
            for (int u = 0; u < kFarBatch / 64; ++u) {
                if (u >= trips) break;
                const bool valid = fids[u] >= 0;
                const unsigned o16 = (unsigned)(valid ? fids[u] : 0) << 4;
                const float4 mc = *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(mac) + o16);
                float4 qa = *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(q0) + o16);
                float4 qb = *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(q1) + o16);
                float x = mc.x - bc0, y = mc.y - bc1, z = mc.z - bc2;
                if (!valid) {
                    x = 1.f;
                    y = 0.f;
                    z = 0.f;
                    qa = make_float4(0.f, 0.f, 0.f, 0.f);
                    qb = qa;
                }
                if (!kQuad) qb = make_float4(0.f, 0.f, 0.f, 0.f);
                const float M = qa.x;
                const float Qxx = kQuad ? qa.y : 0.f, Qyy = kQuad ? qa.z : 0.f, Qzz = kQuad ? qa.w : 0.f;
                // Order 0: softened monopole + quadrupole acceleration at the box centre.
                const float r2s = fmaf(x, x, fmaf(y, y, fmaf(z, z, soft2)));
                const float is = __builtin_amdgcn_rsqf(r2s);
                const float is2 = is * is;
                const float qdx = fmaf(Qxx, x, fmaf(qb.x, y, qb.y * z));
                const float qdy = fmaf(qb.x, x, fmaf(Qyy, y, qb.z * z));
                const float qdz = fmaf(qb.y, x, fmaf(qb.z, y, Qzz * z));
                const float dqd = fmaf(x, qdx, fmaf(y, qdy, z * qdz));
                const float is5 = (is2 * is2) * is;
                const float sc = fmaf(2.5f, dqd * is2, M * r2s);
                c[0] += fmaf(x, sc, -qdx) * is5;
                c[1] += fmaf(y, sc, -qdy) * is5;
                c[2] += fmaf(z, sc, -qdz) * is5;
                // Orders 1-3, unsoftened, in the unit vector n = R/r: order n is
                // stored times h^(n-1).  g2 = 1/r^2, hw = h/r.
                const float r2 = fmaf(x, x, fmaf(y, y, z * z));
                const float w = __builtin_amdgcn_rsqf(r2);
                const float g2 = w * w;
                const float nx = x * w, ny = y * w, nz = z * w;
                const float hw = hg * w;
                const float k2 = hw * g2;      // h/r^3
                const float k3 = M * hw * k2;  // M h^2/r^4
                const float k4 = k3 * hw;      // M h^3/r^5
                // L2 h = (h/r^3) [(M + 5/2 s) I - (3M + 35/2 s) nn + 5 (n p + p n) - Q/r^2],
                // p = Qn/r^2, s = n.p (traceless: xx, yy, xy, xz, yz).
                const float sig = dqd * (g2 * g2);
                const float wg = w * g2;
                const float px = qdx * wg, py = qdy * wg, pz = qdz * wg;
                const float A = -fmaf(17.5f, sig, 3.f * M) * k2;
                const float C = fmaf(2.5f, sig, M) * k2;
                const float B = 5.f * k2;
                const float D = -k2 * g2;
                c[3] += fmaf(A, nx * nx, fmaf(2.f * B, nx * px, fmaf(D, Qxx, C)));
                c[4] += fmaf(A, ny * ny, fmaf(2.f * B, ny * py, fmaf(D, Qyy, C)));
                c[5] += fmaf(A, nx * ny, fmaf(B, fmaf(nx, py, ny * px), D * qb.x));
                c[6] += fmaf(A, nx * nz, fmaf(B, fmaf(nx, pz, nz * px), D * qb.y));
                c[7] += fmaf(A, ny * nz, fmaf(B, fmaf(ny, pz, nz * py), D * qb.z));
                // L3 h^2 = (M h^2/r^4) (15 nnn - 3 sym(n I)): xxx, xxy, xxz, xyy, xyz, yyy, yyz.
                const float e = 15.f * k3, f = 3.f * k3;
                const float xx = nx * nx, yy = ny * ny;
                const float ex = fmaf(e, xx, -f), ey = fmaf(e, yy, -f);
                c[8] += nx * fmaf(-2.f, f, ex);
                c[9] += ny * ex;
                c[10] += nz * ex;
                c[11] += nx * ey;
                c[12] += e * (nx * ny) * nz;
                c[13] += ny * fmaf(-2.f, f, ey);
                c[14] += nz * ey;
                // L4 h^3 = (M h^3/r^5) (-105 nnnn + 15 sym(nn I) - 3 sym(I I)):
                // xxxx, xxxy, xxxz, xxyy, xxyz, xyyy, xyyz, yyyy, yyyz.
                const float a105 = -105.f * k4;
                const float ux = fmaf(a105, xx, 45.f * k4), uy = fmaf(a105, yy, 45.f * k4);
                const float vx = fmaf(a105, xx, 15.f * k4), vy = fmaf(a105, yy, 15.f * k4);
                c[15] += fmaf(xx, fmaf(a105, xx, 90.f * k4), -9.f * k4);
                c[16] += (nx * ny) * ux;
                c[17] += (nx * nz) * ux;
                c[18] += fmaf(xx, fmaf(a105, yy, 15.f * k4), fmaf(15.f * k4, yy, -3.f * k4));
                c[19] += (ny * nz) * vx;
                c[20] += (nx * ny) * uy;
                c[21] += (nx * nz) * vy;
                c[22] += fmaf(yy, fmaf(a105, yy, 90.f * k4), -9.f * k4);
                c[23] += (ny * nz) * uy;
            }
            // Reduce-scatter over the wave: three pair exchanges inside each
            // 16-lane row (row_mirror, row_half_mirror, quad_perm 3210) halve
            // the coefficients a lane holds, 24 -> 12 -> 6 -> 3; the pair i, i^1
            // and the four rows are then summed.  Lane bits 3, 2, 1 select
            // coefficients 12 b3 + 6 b2 + 3 b1 + (0, 1, 2).  Fixed order: the
            // sums are deterministic.
            const bool u8 = lane & 8, u4 = lane & 4, u2 = lane & 2;
            float h12[12], h6[6], h3[3];
#pragma unroll
            for (int k = 0; k < 12; ++k)
                h12[k] = dpp_pair_add<0x140>(u8 ? c[k + 12] : c[k], u8 ? c[k] : c[k + 12]);
#pragma unroll
            for (int k = 0; k < 6; ++k)
                h6[k] = dpp_pair_add<0x141>(u4 ? h12[k + 6] : h12[k], u4 ? h12[k] : h12[k + 6]);
#pragma unroll
            for (int k = 0; k < 3; ++k)
                h3[k] = dpp_pair_add<0x1B>(u2 ? h6[k + 3] : h6[k], u2 ? h6[k] : h6[k + 3]);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                h3[k] = dpp_add<0xB1>(h3[k]);
                h3[k] += __shfl_xor(h3[k], 16, 64);
                h3[k] += __shfl_xor(h3[k], 32, 64);
            }
            if (lane < 16 && !(lane & 1)) {
                const int base = (u8 ? 12 : 0) + (u4 ? 6 : 0) + (u2 ? 3 : 0);
#pragma unroll
                for (int k = 0; k < 3; ++k) lexp[base + k] += h3[k];
            }
            npf += (unsigned long long)cnt;
        };

        // Classifies one candidate node per lane (id < 0: none) against the
        // group.  The node's walk record is read once: an accepted cell goes
        // from registers to the LDS staging list.
        auto process = [&](int id) {
            const bool valid = id >= 0;
            const int idc = valid ? id : 0;  // lanes without a candidate read the root's record
            // 32-bit byte offsets from the uniform array bases (node ids < 2^28).
            const unsigned o16 = (unsigned)idc << 4, o8 = (unsigned)idc << 3;
            const float4 mc = *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(mac) + o16);
            const int2 lk = *reinterpret_cast<const int2*>(reinterpret_cast<const char*>(link) + o8);
            const float4 qa = *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(q0) + o16);
            const float4 qb =
                kQuad ? *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(q1) + o16) : qa;
            const float e0 = fmaxf(fabsf(mc.x - bc0) - bh0, 0.f);
            const float e1 = fmaxf(fabsf(mc.y - bc1) - bh1, 0.f);
            const float e2 = fmaxf(fabsf(mc.z - bc2) - bh2, 0.f);
            const float d2 = fmaf(e0, e0, fmaf(e1, e1, e2 * e2));
            const bool acc = valid && d2 > mc.w;
            const bool leaf = valid && !acc && lk.x < 0;
            const bool open = valid && !acc && lk.x >= 0;
            bool far = false;
            if (kFar && acc) {
                const float rx = mc.x - bc0, ry = mc.y - bc1, rz = mc.z - bc2;
                far = rg2 < rho2 * fmaf(rx, rx, fmaf(ry, ry, rz * rz));
            }
            const bool nacc = acc && !far;  // accepted, evaluated at every target

            const unsigned long long bo = __ballot(open);
            const int no = __popcll(bo);
            if (open) {
                const int pos = nnext + __popcll(bo & lt);
                if (pos < kFl) fl[nb][pos] = lk;
                else if (pos < fcap) nxtg[pos] = lk;
            }
            if (nnext + no > fcap) gov = true;
            nnext = min(nnext + no, fcap);

            // Accepted cells: staged in LDS and evaluated cfull at a time
            // (cfull is a multiple of the subset count, so no subset idles).
            // Cells past a full batch wait in registers during its evaluation.
            const unsigned long long ba = __ballot(nacc);
            const int na = __popcll(ba);
            const int pos = ncl + __popcll(ba & lt);  // < cfull + 63
            if (nacc && pos < cfull) {
                sc0[pos] = mc;
                sc1[pos] = qa;
                if (kQuad) sc2[pos] = qb;
            }
            ncl += na;
            if (ncl >= cfull) {
                flush_cells(cfull);
                if (nacc && pos >= cfull) {
                    sc0[pos - cfull] = mc;
                    sc1[pos - cfull] = qa;
                    if (kQuad) sc2[pos - cfull] = qb;
                }
                ncl -= cfull;
                if (ncl >= cfull) {  // possible when cfull < 64: a second batch
                    flush_cells(cfull);
                    const bool mv = lane < ncl - cfull;
                    float4 v0 = make_float4(0.f, 0.f, 0.f, 0.f), v1 = v0, v2 = v0;
                    if (mv) {
                        v0 = sc0[cfull + lane];
                        v1 = sc1[cfull + lane];
                        if (kQuad) v2 = sc2[cfull + lane];
                    }
                    __syncthreads();
                    if (mv) {
                        sc0[lane] = v0;
                        sc1[lane] = v1;
                        if (kQuad) sc2[lane] = v2;
                    }
                    ncl -= cfull;
                }
            }

            // Far cells: node ids staged kFarBatch at a time; ids past a full
            // batch wait in registers during its evaluation.
            if (kFar) {
                const unsigned long long bf = __ballot(far);
                if (bf) {
                    const int fpos = nfar + __popcll(bf & lt);  // < kFarBatch + 63
                    if (far && fpos < kFarBatch) flist[fpos] = idc;
                    nfar += __popcll(bf);
                    if (nfar >= kFarBatch) {
                        flush_far(kFarBatch);
                        if (far && fpos >= kFarBatch) flist[fpos - kFarBatch] = idc;
                        nfar -= kFarBatch;
                    }
                }
            }

            // Opened leaves: each leaf's lane copies its own particles (a
            // contiguous range of the sorted particles) into the staging list,
            // 64 list slots per round, four loads in flight at a time.
            if (__ballot(leaf) == 0ull) return;  // most candidate batches open no leaf
            const int cnt = leaf ? (lk.y - ~lk.x) : 0;
            const int incl = wave_incl_scan(cnt);
            const int tot = __shfl(incl, 63, 64);
            const int excl = incl - cnt;
            const char* src = reinterpret_cast<const char*>(spos);  // + 32-bit byte offsets
            const unsigned first16 = (unsigned)(leaf ? ~lk.x : 0) << 4;
            for (int t0 = 0; t0 < tot; t0 += 64) {
                // this lane's particles k with t0 <= excl + k < t0 + 64, to
                // slot npl + excl + k - t0 (npl < cfull <= 64 here, so the slot is < 128)
                const int lo = max(0, t0 - excl), hi = min(cnt, t0 + 64 - excl);
                for (int k = lo; k < hi; k += 4) {
                    const unsigned o0 = first16 + ((unsigned)k << 4);
                    const float4 v0 = *reinterpret_cast<const float4*>(src + o0);
                    const float4 v1 = *reinterpret_cast<const float4*>(
                        src + first16 + ((unsigned)min(k + 1, hi - 1) << 4));
                    const float4 v2 = *reinterpret_cast<const float4*>(
                        src + first16 + ((unsigned)min(k + 2, hi - 1) << 4));
                    const float4 v3 = *reinterpret_cast<const float4*>(
                        src + first16 + ((unsigned)min(k + 3, hi - 1) << 4));
                    float4* dst = pl + (npl + excl + k - t0);
                    dst[0] = v0;
                    if (k + 1 < hi) dst[1] = v1;
                    if (k + 2 < hi) dst[2] = v2;
                    if (k + 3 < hi) dst[3] = v3;
                }
                npl += min(64, tot - t0);
                __syncthreads();
                while (npl >= cfull) {  // npl < cfull + 64 <= 128
                    flush_particles(cfull);
                    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
                    const bool mv = lane < npl - cfull;
                    if (mv) v = pl[cfull + lane];
                    __syncthreads();
                    if (mv) pl[lane] = v;
                    npl -= cfull;
                    __syncthreads();
                }
            }
        };

        // Root (node 0), then one tree level per pass; the opened cells' child
        // ranges of a level are its frontier (entries 0..63 in LDS buffer nb).
        process(lane == 0 ? 0 : -1);
        int ncur = nnext;
        __syncthreads();
        while (ncur > 0) {
            maxf = max(maxf, ncur);
            const int cb = nb;
            const int2* curg = cb ? fr1 : fr0;
            nb = cb ^ 1;
            nxtg = nb ? fr1 : fr0;
            nnext = 0;
            for (int base = 0; base < ncur; base += 8) {
                const int ps = base + (lane >> 3);
                const int cs = lane & 7;
                int id = -1;
                if (ps < ncur) {
                    int2 rg;
                    if (base < kFl) {  // base is a multiple of 8: ps < kFl
                        rg = fl[cb][ps];
                    } else {  // (an atomic load keeps the two reads from merging into a flat load)
                        const unsigned long long v = __hip_atomic_load(
                            reinterpret_cast<const unsigned long long*>(curg + ps), __ATOMIC_RELAXED,
                            __HIP_MEMORY_SCOPE_WORKGROUP);
                        rg = make_int2((int)(unsigned)v, (int)(unsigned)(v >> 32));
                    }
                    if (cs < rg.y - rg.x) id = rg.x + cs;
                }
                process(id);
            }
            __syncthreads();
            ncur = nnext;
        }
        if (ncl > 0) flush_cells(ncl);
        if (npl > 0) flush_particles(npl);
        if (kFar && nfar > 0) flush_far(nfar);

        if (nsub > 1) {  // reduce the subsets' partial sums in subset order
            sc0[lane] = make_float4(fxA, fyA, fzA, 0.f);
            sc1[lane] = make_float4(fxB, fyB, fzB, 0.f);
            if (t3) sc2[lane] = make_float4(fxC, fyC, fzC, 0.f);
            __syncthreads();
            if (lead) {
                for (int q = 1; q < nsub; ++q) {
                    const float4 a = sc0[q * nl + tix];
                    const float4 b = sc1[q * nl + tix];
                    fxA += a.x;
                    fyA += a.y;
                    fzA += a.z;
                    fxB += b.x;
                    fyB += b.y;
                    fzB += b.z;
                    if (t3) {
                        const float4 cc = sc2[q * nl + tix];
                        fxC += cc.x;
                        fyC += cc.y;
                        fzC += cc.z;
                    }
                }
            }
            __syncthreads();
        }
        if (kFar && npf > 0) {  // the far cells' expansion at each target
            __syncthreads();      // lane 63's sums are visible
            float L[24];
#pragma unroll
            for (int j = 0; j < 24; ++j) L[j] = lexp[j];
            // Order 1 (zz = -xx - yy), order 2 (xzz, yzz, zzz from the trace
            // conditions), order 3 (the six components with z^2 likewise).
            const float t200 = L[3], t020 = L[4], t110 = L[5], t101 = L[6], t011 = L[7];
            const float t002 = -t200 - t020;
            const float t300 = L[8], t210 = L[9], t201 = L[10], t120 = L[11], t111 = L[12],
                        t030 = L[13], t021 = L[14];
            const float t102 = -t300 - t120, t012 = -t210 - t030, t003 = -t201 - t021;
            const float u400 = L[15], u310 = L[16], u301 = L[17], u220 = L[18], u211 = L[19],
                        u130 = L[20], u121 = L[21], u040 = L[22], u031 = L[23];
            const float u202 = -u400 - u220, u112 = -u310 - u130, u022 = -u220 - u040,
                        u103 = -u301 - u121, u013 = -u211 - u031, u004 = -u202 - u022;
            auto l2p = [&](const float4& me, float& fx, float& fy, float& fz) {
                const float p = (me.x - bc0) * hinv, q = (me.y - bc1) * hinv, s = (me.z - bc2) * hinv;
                const float g2x = fmaf(t200, p, fmaf(t110, q, t101 * s));
                const float g2y = fmaf(t110, p, fmaf(t020, q, t011 * s));
                const float g2z = fmaf(t101, p, fmaf(t011, q, t002 * s));
                const float pp = p * p, qq = q * q, ss = s * s, pq = p * q, ps = p * s, qs = q * s;
                const float g3x = fmaf(t300, pp, fmaf(t120, qq, fmaf(t102, ss,
                                  2.f * fmaf(t210, pq, fmaf(t201, ps, t111 * qs)))));
                const float g3y = fmaf(t210, pp, fmaf(t030, qq, fmaf(t012, ss,
                                  2.f * fmaf(t120, pq, fmaf(t111, ps, t021 * qs)))));
                const float g3z = fmaf(t201, pp, fmaf(t021, qq, fmaf(t003, ss,
                                  2.f * fmaf(t111, pq, fmaf(t102, ps, t012 * qs)))));
                const float p3 = pp * p, q3 = qq * q, s3 = ss * s, pqs = pq * s;
                const float g4x = fmaf(u400, p3, fmaf(u130, q3, fmaf(u103, s3, fmaf(6.f * u211, pqs,
                                  3.f * fmaf(u310, pp * q, fmaf(u301, pp * s, fmaf(u220, p * qq,
                                  fmaf(u121, qq * s, fmaf(u202, p * ss, u112 * (q * ss))))))))));
                const float g4y = fmaf(u310, p3, fmaf(u040, q3, fmaf(u013, s3, fmaf(6.f * u121, pqs,
                                  3.f * fmaf(u220, pp * q, fmaf(u211, pp * s, fmaf(u130, p * qq,
                                  fmaf(u031, qq * s, fmaf(u112, p * ss, u022 * (q * ss))))))))));
                const float g4z = fmaf(u301, p3, fmaf(u031, q3, fmaf(u004, s3, fmaf(6.f * u112, pqs,
                                  3.f * fmaf(u211, pp * q, fmaf(u202, pp * s, fmaf(u121, p * qq,
                                  fmaf(u022, qq * s, fmaf(u103, p * ss, u013 * (q * ss))))))))));
                constexpr float kSixth = 1.f / 6.f;
                fx += fmaf(-kSixth, g4x, fmaf(0.5f, g3x, L[0] - g2x));
                fy += fmaf(-kSixth, g4y, fmaf(0.5f, g3y, L[1] - g2y));
                fz += fmaf(-kSixth, g4z, fmaf(0.5f, g3z, L[2] - g2z));
            };
            if (lead) {
                l2p(meA, fxA, fyA, fzA);
                if (hasB) l2p(meB, fxB, fyB, fzB);
                if (hasC) l2p(meC, fxC, fyC, fzC);
            }
        }
        // A group whose frontier overflowed skipped cells: its accelerations
        // are NaN (error bit 1), never silently incomplete.
        ovf = ovf || gov;
        const double gs = gov ? __builtin_nan("") : G;
        if (lead) {
            const unsigned j = perm[kA];
            ax[j] = gs * (double)fxA;
            ay[j] = gs * (double)fyA;
            az[j] = gs * (double)fzA;
            if (hasB) {
                const unsigned jb = perm[kB];
                ax[jb] = gs * (double)fxB;
                ay[jb] = gs * (double)fyB;
                az[jb] = gs * (double)fzB;
            }
            if (hasC) {
                const unsigned jc = perm[kC];
                ax[jc] = gs * (double)fxC;
                ay[jc] = gs * (double)fyC;
                az[jc] = gs * (double)fzC;
            }
        }
        tot_pp += npp * (unsigned long long)nact;
        tot_pc += npc * (unsigned long long)nact;
        tot_pf += npf;
        tot_pft += npf * (unsigned long long)nact;
      }
    }
    if (lane == 0) {
        atomicAdd(&ctr->pp, tot_pp);
        atomicAdd(&ctr->pc, tot_pc);
        if (kFar) {
            atomicAdd(&ctr->pf, tot_pf);
            atomicAdd(&ctr->pft, tot_pft);
        }
        atomicMax(&ctr->max_frontier, maxf);
        if (ovf) raise_flag(ctr, 1);
    }
}

// ------------------------------------------------------------------ host
// Node pool size: an upper bound on the tree's nodes (every particle a leaf
// on a chain of up to 21 levels per cap + 1 particles), never above
// kMaxNodes: the walk addresses node records with 32-bit byte offsets
// (16-byte records), so node ids must stay below 2^28.  A tree with more
// nodes than the pool is reported (error bit 2), never walked.
// NB_BH_NODE_CAP (testing) can only lower it, to exercise that report.
int64_t node_capacity(int64_t n, int cap) {
    if (n <= 0) return 0;
    cap = std::max(1, cap);
    const int64_t bound =
        std::min<int64_t>(kMaxNodes, n + (int64_t)kMaxDigits * ((n + cap) / (cap + 1)) + 1);
    if (const char* e = std::getenv("NB_BH_NODE_CAP"))
        return std::max<int64_t>(1, std::min<int64_t>(bound, std::atoll(e)));
    return bound;
}

// Sorted particles per thread of the build kernels (1, 2 or 4) for n keys.
// NB_BUILD_ROWS (tuning) forces it.
int build_rows(int64_t n) {
    if (const char* e = std::getenv("NB_BUILD_ROWS")) {
        const int r = std::atoi(e);
        if (r == 1 || r == 2 || r == 4) return r;
    }
    return n <= ((int64_t)1 << 21) ? 2 : kBuildRowsMax;
}

// Keys per radix-pass tile for n keys (4096 or 8192) and earlier tiles read
// per look-back round trip (8, 16 or 32).  NB_SORT_TILE and NB_SORT_WINDOW
// (tuning) force them.
int sort_tile(int64_t n) {
    if (const char* e = std::getenv("NB_SORT_TILE")) {
        const int t = std::atoi(e);
        if (t == 4096 || t == 8192) return t;
    }
    return 4096;
}

int sort_window(int64_t n) {
    if (const char* e = std::getenv("NB_SORT_WINDOW")) {
        const int t = std::atoi(e);
        if (t == 8 || t == 16 || t == 32) return t;
    }
    return 16;
}

// Walk frontier capacity per block for opening angle theta.  NB_WALK_FCAP
// (testing) sets it directly, e.g. small enough to exercise the overflow
// report (error bit 1); kernels stay in bounds at any value >= 1.
int walk_fcap(double theta) {
    if (const char* e = std::getenv("NB_WALK_FCAP")) return std::max(1, std::min(kWalkFcapMax, std::atoi(e)));
    if (!(theta < 0.5)) return kWalkFcap;
    const double r = 0.5 / std::max(theta, 1e-3);
    return (int)std::min<double>(kWalkFcapMax, std::ceil(kWalkFcap * r * r * r / 256.0) * 256.0);
}

// Far-cell ratio of the quadrupole walk: a cell accepted for a group goes
// into the group's local expansion when the group box's half diagonal is
// below ratio * min(theta, 0.5) times the distance from the box centre to the
// cell's centre of mass (the expansion's fourth-order remainder then stays
// below the quadrupole's own truncation error).  NB_WALK_FAR (tuning) sets
// the ratio; 0 evaluates every accepted cell at every target.
double walk_far_ratio() {
    if (const char* e = std::getenv("NB_WALK_FAR")) return std::max(0.0, std::min(1.0, std::atof(e)));
    return kWalkFarRatio;
}

int64_t walk_scratch_bytes(double theta, int cus) {
    int bpc = kWalkBlocksPerCU;
    if (const char* e = std::getenv("NB_WALK_BPC")) bpc = std::max(1, std::min(64, std::atoi(e)));
    return (int64_t)std::max(1, cus) * bpc * 2 * walk_fcap(theta) * (int64_t)sizeof(int2);
}

int64_t bh_workspace_bytes(int64_t n, int leaf_cap, double theta, int cus) {
    if (n <= 0) return 0;
    const int64_t nblk = (n + kBlk - 1) / kBlk, nsort = (n + sort_tile(n) - 1) / sort_tile(n);
    const int64_t ncap = node_capacity(n, leaf_cap);
    const int64_t per_particle = 8 + 8 + 4 + 4 + 16 + 16 + 1 + 1 + 1 + 4 + 1 + 4 + 4 + 4;
    const int64_t per_node = 4 * 4 + 1 + 16 + 8 + 16 + 16;
    return n * per_particle + 4 + ncap * per_node + std::min<int64_t>(ncap, kMomNodes) * 80 +
           (nsort + (nsort + 7) / 8) * kSortBins * kSortPasses * 4 +
           (int64_t)sizeof(SortState) + nblk * kCols * 4 +
           (int64_t)std::max(kSortBins, kCols) * 4 + 1024 * 6 * 8 + 4096 +
           walk_scratch_bytes(theta, cus);
}

int bh_create(BHWorkspace** w) {
    *w = new BHWorkspace();
    return NB_OK;
}

void bh_destroy(BHWorkspace* w) { delete w; }

int bh_eval(BHWorkspace* w, double G, double soft2, double theta, int leaf_cap, int multipole,
            int64_t n64, const double* m, const double* x, const double* y, const double* z,
            double* ax, double* ay, double* az, hipStream_t s, const hipEvent_t* ev) {
    if (n64 <= 0) return NB_OK;
    if (n64 > (int64_t)1 << 28) {
        set_error("Barnes-Hut: n=%lld exceeds 2^28", (long long)n64);
        return NB_ERR_INVALID;
    }
    const int n = (int)n64;
    const int nblk = grid_for(n, kBlk);
    const int rows = build_rows(n);
    const int ntile = grid_for(n, kBlk * rows);  // build blocks
    const int tile = sort_tile(n);
    const int win = sort_window(n);
    const int nsort = grid_for(n, tile);
    const int nbb = std::min(512, nblk);  // partial boxes (every key block reduces them)
    const int64_t ncap = node_capacity(n, leaf_cap);  // <= kMaxNodes

    NB_TRY(w->root.reserve(sizeof(RootInfo)));
    NB_TRY(w->levels.reserve(sizeof(LevelInfo)));
    {
        const bool fresh = w->counters.ptr == nullptr;
        NB_TRY(w->counters.reserve(sizeof(Counters)));
        // The sticky flag word outlives evaluations: it starts at zero.
        if (fresh) NB_HIP(hipMemsetAsync(w->counters.ptr, 0, sizeof(Counters), s));
    }
    NB_TRY(w->bbox_part.reserve((size_t)nbb * 6 * sizeof(double)));
    NB_TRY(w->keys0.reserve((size_t)n * 8));
    NB_TRY(w->keys1.reserve((size_t)n * 8));
    NB_TRY(w->vals0.reserve((size_t)n * 4));
    NB_TRY(w->vals1.reserve((size_t)n * 4));
    const int ngrp = (nsort + win - 1) / win;  // tile groups of a resident pass
    // look-back words, then the tile groups' totals
    NB_TRY(w->hist.reserve((size_t)(nsort + ngrp) * kSortBins * kSortPasses * 4));
    NB_TRY(w->sortst.reserve(sizeof(SortState)));
    NB_TRY(w->totals.reserve((size_t)std::max(kSortBins, kCols) * 4));
    NB_TRY(w->spos.reserve((size_t)n * 16));
    NB_TRY(w->upos.reserve((size_t)n * 16));
    NB_TRY(w->lvl_leaf.reserve((size_t)n));
    NB_TRY(w->lvl_sh.reserve((size_t)n));
    NB_TRY(w->blkcnt.reserve((size_t)ntile * kCols * 4));
    NB_TRY(w->gflag.reserve((size_t)n));
    NB_TRY(w->gstart.reserve((size_t)(n + 1) * 4));
    NB_TRY(w->gorder.reserve((size_t)n * 4));
    NB_TRY(w->glev.reserve((size_t)n));
    NB_TRY(w->leaf_of.reserve((size_t)n * 4));
    NB_TRY(w->n_pstart.reserve((size_t)ncap * 4));
    NB_TRY(w->n_pend.reserve((size_t)ncap * 4));
    NB_TRY(w->n_child.reserve((size_t)ncap * 4));
    NB_TRY(w->n_cend.reserve((size_t)ncap * 4));
    NB_TRY(w->n_level.reserve((size_t)ncap));
    NB_TRY(w->n_mom.reserve((size_t)std::min<int64_t>(ncap, kMomNodes) * 10 * 8));
    NB_TRY(w->leaves.reserve((size_t)n * 4));
    NB_TRY(w->w_mac.reserve((size_t)ncap * 16));
    NB_TRY(w->w_link.reserve((size_t)ncap * 8));
    NB_TRY(w->w_q0.reserve((size_t)ncap * 16));
    NB_TRY(w->w_q1.reserve((size_t)ncap * 16));

    RootInfo* root = w->root.as<RootInfo>();
    LevelInfo* li = w->levels.as<LevelInfo>();
    Counters* ctr = w->counters.as<Counters>();
    // The key kernel writes the buffer pair that makes the last radix pass
    // write keys0/vals0 (the passes alternate between the pairs).
    constexpr bool kOddPasses = kSortPasses % 2 == 1;
    auto* k0 = (kOddPasses ? w->keys1 : w->keys0).as<unsigned long long>();
    auto* k1 = (kOddPasses ? w->keys0 : w->keys1).as<unsigned long long>();
    auto* v0 = (kOddPasses ? w->vals1 : w->vals0).as<unsigned>();
    auto* v1 = (kOddPasses ? w->vals0 : w->vals1).as<unsigned>();

    SortState* ss = w->sortst.as<SortState>();
    hipLaunchKernelGGL(k_bbox_partial, dim3(nbb), dim3(kBlk), 0, s, n, x, y, z,
                       w->bbox_part.as<double>(), ctr, ss);
    if (ev && ev[1]) NB_HIP(hipEventRecord(ev[1], s));
    unsigned* look = w->hist.as<unsigned>();
    const int nlook = (nsort + ngrp) * kSortBins * kSortPasses;  // zeroed by the key kernel
    // Key grid: fewer blocks for small n (each block adds its digit table
    // to the global counts once); NB_KEY_BLOCKS (tuning) overrides.
    int kblocks = n <= (1 << 22) ? kKeyBlocks / 2 : kKeyBlocks;
    if (const char* e = std::getenv("NB_KEY_BLOCKS")) kblocks = std::max(1, std::min(4096, std::atoi(e)));
    hipLaunchKernelGGL(k_keys, dim3(std::min(nblk, kblocks)), dim3(kBlk), 0, s, n, x, y, z, m,
                       nbb, w->bbox_part.as<double>(), root, k0, v0, w->upos.as<float4>(), ctr, ss,
                       look, nlook);
    NB_LAUNCH_CHECK();
    if (ev && ev[2]) NB_HIP(hipEventRecord(ev[2], s));
    if (w->sort_slots_key != tile * 64 + win) {
        // Radix-pass blocks the whole GPU holds at once for this tile shape.
        int dev = 0, cus = 0, per_cu = 0;
        NB_HIP(hipGetDevice(&dev));
        NB_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
#define NB_OCC(T, R, W) \
    NB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_onesweep<T, R, W>, T, 0))
        switch (tile * 64 + win) {
            case 4096 * 64 + 16: NB_OCC(512, 8, 16); break;
            case 4096 * 64 + 32: NB_OCC(512, 8, 32); break;
            case 8192 * 64 + 16: NB_OCC(512, 16, 16); break;
            case 8192 * 64 + 32: NB_OCC(512, 16, 32); break;
            case 8192 * 64 + 8: NB_OCC(512, 16, 8); break;
            default: NB_OCC(512, 8, 16); break;
        }
#undef NB_OCC
        w->sort_slots = cus * per_cu;
        if (const char* e = std::getenv("NB_SORT_TICKET")) w->sort_slots = 0;  // testing: always the ticket
        w->sort_slots_key = tile * 64 + win;
    }
    for (int p = 0; p < kSortPasses; ++p) {
        unsigned* lp = look + (size_t)p * nsort * kSortBins;
        unsigned* gp = look + (size_t)kSortPasses * nsort * kSortBins + (size_t)p * ngrp * kSortBins;
        float4* gath = p == kSortPasses - 1 ? w->spos.as<float4>() : nullptr;
        const int resident = nsort <= w->sort_slots / 2 ? 1 : 0;  // half: margin for other work
#define NB_PASS(T, R, W)                                                                         \
    hipLaunchKernelGGL((k_onesweep<T, R, W>), dim3(nsort), dim3(T), 0, s, n, k0, v0, k1, v1, p, \
                       ss, lp, ctr, w->upos.as<float4>(), gath, resident, resident ? gp : nullptr)
        switch (tile * 64 + win) {
            case 4096 * 64 + 16: NB_PASS(512, 8, 16); break;
            case 4096 * 64 + 32: NB_PASS(512, 8, 32); break;
            case 8192 * 64 + 16: NB_PASS(512, 16, 16); break;
            case 8192 * 64 + 32: NB_PASS(512, 16, 32); break;
            case 8192 * 64 + 8: NB_PASS(512, 16, 8); break;
            default: NB_PASS(512, 8, 16); break;
        }
#undef NB_PASS
        std::swap(k0, k1);
        std::swap(v0, v1);
    }
    NB_LAUNCH_CHECK();
    // The sorted keys and permutation are in keys0/vals0 (k0/v0).
    if (ev && ev[3]) NB_HIP(hipEventRecord(ev[3], s));
    // (The gather ran in the last radix pass: the gather stage is empty.)
    if (ev && ev[4]) NB_HIP(hipEventRecord(ev[4], s));
    if (w->group_size == 0) {  // walk group size: NB_WALK_GROUP overrides the default
        w->group_size = kGroupDefault;
        if (const char* e = std::getenv("NB_WALK_GROUP"))
            w->group_size = std::max(1, std::min(kGroup, std::atoi(e)));
    }
#define NB_LEVELS(R)                                                                          \
    hipLaunchKernelGGL(k_levels<R>, dim3(ntile), dim3(kBlk), 0, s, n, k0, leaf_cap,             \
                       w->group_size, root, w->lvl_leaf.as<signed char>(),                       \
                       w->lvl_sh.as<signed char>(), w->gflag.as<unsigned char>(),                \
                       w->glev.as<signed char>(), ntile, w->blkcnt.as<int>(), ctr)
    if (rows == 4) NB_LEVELS(4);
    else if (rows == 2) NB_LEVELS(2);
    else NB_LEVELS(1);
#undef NB_LEVELS
    hipLaunchKernelGGL(k_col_scan, dim3(kCols), dim3(1024), 0, s, ntile, w->blkcnt.as<int>(),
                       w->totals.as<int>());
#define NB_EMIT(R)                                                                            \
    hipLaunchKernelGGL(k_emit<R>, dim3(ntile), dim3(kBlk), 0, s, n,                             \
                       w->lvl_leaf.as<signed char>(), w->lvl_sh.as<signed char>(), ntile,        \
                       w->blkcnt.as<int>(), w->totals.as<int>(), li, (int)ncap,                  \
                       w->n_pstart.as<int>(), w->n_pend.as<int>(), w->n_child.as<int>(),         \
                       w->n_cend.as<int>(), w->n_level.as<signed char>(), w->leaf_of.as<int>(),  \
                       w->leaves.as<int>(), w->gflag.as<unsigned char>(), w->gstart.as<int>(),   \
                       w->gorder.as<int>(), ctr)
    if (rows == 4) NB_EMIT(4);
    else if (rows == 2) NB_EMIT(2);
    else NB_EMIT(1);
#undef NB_EMIT
    NB_LAUNCH_CHECK();
    if (ev && ev[5]) NB_HIP(hipEventRecord(ev[5], s));
    const double inv_theta2 = theta > 0.0 ? 1.0 / (theta * theta) : INFINITY;
    hipLaunchKernelGGL(k_leaf_moments, dim3(grid_for(n, kBlk)), dim3(kBlk), 0, s, li, root,
                       w->n_pstart.as<int>(), w->n_pend.as<int>(), w->n_child.as<int>(),
                       w->n_cend.as<int>(), w->n_level.as<signed char>(), w->spos.as<float4>(),
                       w->n_mom.as<double>(), w->w_mac.as<float4>(), w->w_link.as<int2>(),
                       w->w_q0.as<float4>(), w->w_q1.as<float4>(), inv_theta2,
                       multipole == NB_QUADRUPOLE ? 1 : 0, n, w->leaves.as<int>(), ctr);
    // Up-sweep schedule from the level sizes of a recent evaluation: a level
    // with at least kDenseLevel nodes gets its own launch, a run of sparser
    // levels one single-block launch.  Without sizes every level gets a launch.
    bool capturing = false;
    {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        NB_HIP(hipStreamIsCapturing(s, &cs));
        capturing = cs != hipStreamCaptureStatusNone;
    }
    if (w->hint_pending && !capturing) {
        const hipError_t q = hipEventQuery(w->hint_done);
        if (q == hipSuccess) {
            std::memcpy(w->hint, w->hint_host, sizeof(w->hint));
            w->have_hint = true;
            w->hint_pending = false;
        } else if (q == hipErrorNotReady) {
            (void)hipGetLastError();  // not an error: the copy is still in flight
        } else {
            NB_HIP(q);
        }
    }
    // NB_UPSWEEP_PER_LEVEL (testing) keeps one launch per level;
    // NB_UPSWEEP_DENSE (tuning) sets the level size of a launch of its own.
    const bool planned = w->have_hint && !capturing && !std::getenv("NB_UPSWEEP_PER_LEVEL");
    int dense = kDenseLevel;
    if (const char* e = std::getenv("NB_UPSWEEP_DENSE")) dense = std::max(1, std::atoi(e));
    for (int l = kLevels - 1; l >= 0;) {
        if (!planned || w->hint[l] >= dense) {
            // One thread per node: a level has at most min(8^l, n) nodes (or
            // about as many as it had, when planned).  The grid is capped
            // (the kernel strides): deep levels are mostly empty.
            int64_t lvl_max = l < 8 ? std::min<int64_t>((int64_t)1 << (3 * l), n) : n;
            if (planned) lvl_max = std::min<int64_t>(lvl_max, w->hint[l] + w->hint[l] / 4 + kBlk);
            const int ugrid = (int)std::min<int64_t>(
                kUpsweepBlocks, std::max<int64_t>(1, (lvl_max + kBlk - 1) / kBlk));
            hipLaunchKernelGGL(k_upsweep, dim3(ugrid), dim3(kBlk), 0, s, l, li, root,
                               w->n_pstart.as<int>(), w->n_pend.as<int>(), w->n_child.as<int>(),
                               w->n_cend.as<int>(), w->n_level.as<signed char>(),
                               w->spos.as<float4>(), w->n_mom.as<double>(), w->w_mac.as<float4>(),
                               w->w_link.as<int2>(), w->w_q0.as<float4>(), w->w_q1.as<float4>(),
                               inv_theta2, multipole == NB_QUADRUPOLE ? 1 : 0, n, ctr);
            --l;
        } else {
            int lo = l;
            while (lo > 0 && w->hint[lo - 1] < dense) --lo;
            hipLaunchKernelGGL(k_upsweep_band, dim3(1), dim3(kBandThreads), 0, s, l, lo, li, root,
                               w->n_pstart.as<int>(), w->n_pend.as<int>(), w->n_child.as<int>(),
                               w->n_cend.as<int>(), w->n_level.as<signed char>(),
                               w->spos.as<float4>(), w->n_mom.as<double>(), w->w_mac.as<float4>(),
                               w->w_link.as<int2>(), w->w_q0.as<float4>(), w->w_q1.as<float4>(),
                               inv_theta2, multipole == NB_QUADRUPOLE ? 1 : 0, n, ctr);
            l = lo - 1;
        }
    }
    NB_LAUNCH_CHECK();
    if (!w->hint_pending && !capturing) {
        // Level sizes of this evaluation to pinned host memory, off the main
        // stream (the walk does not wait for the copy).
        if (!w->hint_stream) {
            NB_HIP(hipStreamCreateWithFlags(&w->hint_stream, hipStreamNonBlocking));
            NB_HIP(hipEventCreateWithFlags(&w->hint_ready, hipEventDisableTiming));
            NB_HIP(hipEventCreateWithFlags(&w->hint_done, hipEventDisableTiming));
            NB_HIP(hipHostMalloc((void**)&w->hint_host, sizeof(w->hint), hipHostMallocDefault));
        }
        NB_HIP(hipEventRecord(w->hint_ready, s));
        NB_HIP(hipStreamWaitEvent(w->hint_stream, w->hint_ready, 0));
        NB_HIP(hipMemcpyAsync(w->hint_host,
                              reinterpret_cast<const char*>(li) + offsetof(LevelInfo, count),
                              sizeof(w->hint), hipMemcpyDeviceToHost, w->hint_stream));
        NB_HIP(hipEventRecord(w->hint_done, w->hint_stream));
        w->hint_pending = true;
    }
    if (ev && ev[6]) NB_HIP(hipEventRecord(ev[6], s));

    const int ngroups = grid_for(n, 16);  // upper bound on walk groups in flight
    const bool quad = multipole == NB_QUADRUPOLE;
    const double rho = walk_far_ratio() * std::min(theta, 0.5);  // rho <= 0.2 by default
    const float rho2 = (float)(rho * rho);
    const bool far = quad && rho2 > 0.f;
    if (w->cus == 0) {
        int dev = 0;
        NB_HIP(hipGetDevice(&dev));
        NB_HIP(hipDeviceGetAttribute(&w->cus, hipDeviceAttributeMultiprocessorCount, dev));
        w->cus = std::max(1, w->cus);
    }
    int bpc = far ? kWalkBlocksPerCUFar : kWalkBlocksPerCU;
    if (const char* e = std::getenv("NB_WALK_BPC")) bpc = std::max(1, std::min(64, std::atoi(e)));
    // The frontier scratch serves the largest grid either walk variant uses.
    w->walk_blocks = std::max(w->walk_blocks, w->cus * std::max(bpc, kWalkBlocksPerCU));
    const int wblocks = std::min(ngroups, w->cus * bpc);
    const int fcap = walk_fcap(theta);
    NB_TRY(w->walk_scratch.reserve((size_t)w->walk_blocks * 2 * fcap * sizeof(int2)));
    const float s2 = (float)soft2;
#define NB_WALK(Q, E, F)                                                                       \
    hipLaunchKernelGGL((k_walk<Q, E, F>), dim3(wblocks), dim3(64), 0, s, n,                      \
                       w->spos.as<float4>(), v0, w->w_mac.as<float4>(), w->w_link.as<int2>(),   \
                       w->w_q0.as<float4>(), w->w_q1.as<float4>(), s2, G, ax, ay, az,          \
                       w->walk_scratch.as<int2>(), fcap, w->gstart.as<int>(),                \
                       w->gorder.as<int>(), li, ctr, rho2)
    if (quad) {
        if (s2 > 0.f) {
            if (far) NB_WALK(true, false, true);
            else NB_WALK(true, false, false);
        } else {
            if (far) NB_WALK(true, true, true);
            else NB_WALK(true, true, false);
        }
    } else {
        if (s2 > 0.f) NB_WALK(false, false, false);
        else NB_WALK(false, true, false);
    }
#undef NB_WALK
    NB_LAUNCH_CHECK();
    if (ev && ev[7]) NB_HIP(hipEventRecord(ev[7], s));
    w->last_n = n;
    w->last_cap = leaf_cap;
    w->last_valid = 1;
    return NB_OK;
}

int bh_stats(BHWorkspace* w, hipStream_t s, nb_bh_stats* out) {
    if (!w->last_valid) {
        set_error("Barnes-Hut stats: no evaluation yet");
        return NB_ERR_STATE;
    }
    NB_HIP(hipStreamSynchronize(s));
    RootInfo r;
    LevelInfo li;
    Counters c;
    NB_HIP(hipMemcpy(&r, w->root.ptr, sizeof(r), hipMemcpyDeviceToHost));
    NB_HIP(hipMemcpy(&li, w->levels.ptr, sizeof(li), hipMemcpyDeviceToHost));
    NB_HIP(hipMemcpy(&c, w->counters.ptr, sizeof(c), hipMemcpyDeviceToHost));
    out->n_nodes = li.total;
    out->n_leaves = c.n_leaves;
    out->max_level = li.max_level;
    out->level_limit = r.lmax;
    out->oversize_cells = c.oversize;
    out->pp_interactions = (int64_t)c.pp;
    out->pc_interactions = (int64_t)c.pc;
    out->far_cells = (int64_t)c.pf;
    out->far_target_terms = (int64_t)c.pft;
    out->max_frontier = c.max_frontier;
    out->frontier_overflow = c.overflow;
    for (int a = 0; a < 3; ++a) out->root_center[a] = r.center[a];
    out->root_size = r.size;
    out->n_groups = li.n_groups;
    const int all = c.overflow | c.sticky;
    out->flags_since_read = all;
    out->leaf_capacity = w->last_cap;
    if (c.sticky) {  // reading the flags clears the sticky word
        const int zero = 0;
        NB_HIP(hipMemcpy(reinterpret_cast<char*>(w->counters.ptr) + offsetof(Counters, sticky),
                         &zero, sizeof(zero), hipMemcpyHostToDevice));
    }
    if (all) {
        set_error("Barnes-Hut: invalid evaluation (last evaluation's flags %d, flags since the "
                  "last read %d: 1=walk frontier, 2=node pool, 4=tree consistency, 8=sort "
                  "look-back wait, 16=non-finite position or mass)",
                  c.overflow, all);
        // A capacity or consistency failure NaNs accelerations, which can reach
        // later evaluations as non-finite input: it is the cause to report.
        return (all & ~kInputBad) ? NB_ERR_CAPACITY : NB_ERR_INVALID;
    }
    return NB_OK;
}

// ORs the current evaluation's error flags into *dst (stream-ordered after
// the evaluation), so a caller that shares one calculator between several
// systems can attribute a failure to the system that caused it.
__global__ void k_flags_to(const Counters* __restrict__ ctr, int* __restrict__ dst) {
    if (threadIdx.x == 0) *dst |= ctr->overflow;
}

int bh_flags_to(BHWorkspace* w, int* dst, hipStream_t s) {
    if (!w->counters.ptr) return NB_OK;
    hipLaunchKernelGGL(k_flags_to, dim3(1), dim3(64), 0, s, w->counters.as<Counters>(), dst);
    NB_LAUNCH_CHECK();
    return NB_OK;
}

int bh_topology(BHWorkspace* w, hipStream_t s, int64_t n, int32_t* depth, uint64_t* path,
                uint64_t* keys, uint32_t* perm) {
    if (!w->last_valid || n != w->last_n) {
        set_error("Barnes-Hut topology: n=%lld does not match the last evaluation (%lld)",
                  (long long)n, (long long)w->last_n);
        return NB_ERR_STATE;
    }
    NB_HIP(hipStreamSynchronize(s));
    LevelInfo li;
    NB_HIP(hipMemcpy(&li, w->levels.ptr, sizeof(li), hipMemcpyDeviceToHost));
    if ((size_t)li.total > w->n_level.bytes) {  // the node pool overflowed (error bit 2)
        set_error("Barnes-Hut topology: the tree (%d nodes) exceeded the node pool", li.total);
        return NB_ERR_CAPACITY;
    }
    std::vector<unsigned long long> k(n);
    std::vector<unsigned> p(n);
    std::vector<int> lf(n);
    std::vector<signed char> lv(li.total);
    NB_HIP(hipMemcpy(k.data(), w->keys0.ptr, n * 8, hipMemcpyDeviceToHost));
    NB_HIP(hipMemcpy(p.data(), w->vals0.ptr, n * 4, hipMemcpyDeviceToHost));
    NB_HIP(hipMemcpy(lf.data(), w->leaf_of.ptr, n * 4, hipMemcpyDeviceToHost));
    NB_HIP(hipMemcpy(lv.data(), w->n_level.ptr, li.total, hipMemcpyDeviceToHost));
    for (int64_t q = 0; q < n; ++q) {
        const unsigned o = p[q];
        if (o >= (unsigned)n || lf[q] < 0 || lf[q] >= li.total) {
            set_error("Barnes-Hut topology: malformed tree (particle %lld)", (long long)q);
            return NB_ERR_CAPACITY;
        }
        const int L = lv[lf[q]];
        if (depth) depth[o] = L;
        if (path) path[o] = L == 0 ? 0ull : (k[q] >> (63 - 3 * L));
        if (keys) keys[o] = k[q];
        if (perm) perm[q] = o;
    }
    return NB_OK;
}

int bh_nodes(BHWorkspace* w, hipStream_t s, int64_t capacity, int64_t* count, int32_t* level,
             uint64_t* path, int32_t* first, int32_t* last, double* mass, double* com,
             double* quad) {
    if (!w->last_valid) {
        set_error("Barnes-Hut nodes: no evaluation yet");
        return NB_ERR_STATE;
    }
    NB_HIP(hipStreamSynchronize(s));
    LevelInfo li;
    NB_HIP(hipMemcpy(&li, w->levels.ptr, sizeof(li), hipMemcpyDeviceToHost));
    if ((size_t)li.total > w->n_level.bytes) {  // the node pool overflowed (error bit 2)
        set_error("Barnes-Hut nodes: the tree (%d nodes) exceeded the node pool", li.total);
        return NB_ERR_CAPACITY;
    }
    *count = li.total;
    if (capacity < li.total) return NB_OK;  // caller asked for the count only
    const int64_t nn = li.total;
    // Nodes of the top kMomLevels levels (numbered first) have FP64 records;
    // the others are read back from their FP32 walk records.
    const int64_t n64 = std::min<int64_t>(nn, li.start[kMomLevels]);
    std::vector<signed char> lv(nn);
    std::vector<int> ps(nn), pe(nn);
    std::vector<double> mo((size_t)nn * 10);
    std::vector<float4> fm(nn), fa(nn), fb(nn);
    std::vector<unsigned long long> k(w->last_n);
    NB_HIP(hipMemcpy(lv.data(), w->n_level.ptr, nn, hipMemcpyDeviceToHost));
    NB_HIP(hipMemcpy(ps.data(), w->n_pstart.ptr, nn * 4, hipMemcpyDeviceToHost));
    NB_HIP(hipMemcpy(pe.data(), w->n_pend.ptr, nn * 4, hipMemcpyDeviceToHost));
    NB_HIP(hipMemcpy(mo.data(), w->n_mom.ptr, n64 * 80, hipMemcpyDeviceToHost));
    NB_HIP(hipMemcpy(fm.data(), w->w_mac.ptr, nn * 16, hipMemcpyDeviceToHost));
    NB_HIP(hipMemcpy(fa.data(), w->w_q0.ptr, nn * 16, hipMemcpyDeviceToHost));
    NB_HIP(hipMemcpy(fb.data(), w->w_q1.ptr, nn * 16, hipMemcpyDeviceToHost));
    NB_HIP(hipMemcpy(k.data(), w->keys0.ptr, w->last_n * 8, hipMemcpyDeviceToHost));
    for (int64_t q = n64; q < nn; ++q) {
        double* r = &mo[(size_t)q * 10];
        const double v[10] = {fa[q].x, fm[q].x, fm[q].y, fm[q].z, fa[q].y,
                              fa[q].z, fa[q].w, fb[q].x, fb[q].y, fb[q].z};
        for (int j = 0; j < 10; ++j) r[j] = v[j];
    }
    for (int64_t q = 0; q < nn; ++q) {
        if (ps[q] < 0 || ps[q] >= w->last_n) {
            set_error("Barnes-Hut nodes: malformed tree (node %lld)", (long long)q);
            return NB_ERR_CAPACITY;
        }
        const int L = lv[q];
        if (level) level[q] = L;
        if (path) path[q] = L == 0 ? 0ull : (k[ps[q]] >> (63 - 3 * L));
        if (first) first[q] = ps[q];
        if (last) last[q] = pe[q];
        if (mass) mass[q] = mo[q * 10];
        if (com)
            for (int a = 0; a < 3; ++a) com[q * 3 + a] = mo[q * 10 + 1 + a];
        if (quad)
            for (int a = 0; a < 6; ++a) quad[q * 6 + a] = mo[q * 10 + 4 + a];
    }
    return NB_OK;
}

}  // namespace nb

extern "C" int64_t nb_bh_node_capacity(int64_t n, int leaf_capacity) {
    return nb::node_capacity(n, leaf_capacity);
}
