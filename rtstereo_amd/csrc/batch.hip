// Batched execution of stereo pairs of different sizes under one pipeline
// configuration (rank 7x7 -> AD, LoG -> AD and NCC combined in one pass,
// box aggregation, WTA + right view + left-right check).
//
// The pairs are cut into chunks whose device footprint fits the byte budget.
// Within a chunk the stages run stage by stage over all of its pairs:
//   per pair   gray, rank and LoG images, then the fused three-term cost
//              (disparities per block chosen by the tuner);
//   one launch box aggregation of every pair's volume by a persistent kernel
//              that pulls (pair, disparity, row strip, column tile) work
//              items from a device counter (strip height chosen by the
//              tuner), so pairs of all sizes share one grid;
//   per pair   WTA with the right view and the check.
// The chunk's whole launch sequence is captured into a HIP graph the first
// time the same list of pairs (sizes and buffers) is run and replayed after.
//
// Any other configuration (adaptive weights in FP16 or FP32, census +
// cross-based aggregation, ...) runs through resident pipelines, one per
// pair shape (W, H, D): chunks are cut so that the pipelines of a chunk's
// shapes fit the budget, and a chunk's whole launch sequence (every pair
// through its shape's pipeline, in order) is captured into one HIP graph
// and replayed like the specialised path's.
//
// The reference has no counterpart: its driver runs one pair per process
// (seq/Main.cpp:47-299).
#include "rts_internal.h"

#include <string.h>

#include <algorithm>
#include <map>
#include <memory>
#include <set>
#include <tuple>
#include <vector>

namespace {

using namespace rts;

constexpr int BQ_RMAX = 8;   // box radius limit of the persistent kernel

struct BoxPair {
    const float *in;
    float *out;
    int W, H, D, nx, ny;
    unsigned tile0;   // first work item of this pair
};

// Work item t of pair p: d = t / (ny*nx), strip = (t / nx) % ny, tile = t % nx.
constexpr int BQ_MAXP = 256;   // pairs per chunk

// A work item's position: its pair's buffers and sizes, tile origin, plane.
struct BoxItem {
    const float *src;
    float *dst;
    int W, H, x0, y0;
    bool inner;   // the tile's window lies inside the image (no wrap)
};

// A work item is a TX x TY tile of one plane (TX = 64 or 128 columns): its
// (TY + 2R) x (TX + 2R) window is staged in LDS, so the pair (TX, TY) sets
// the block's LDS footprint (with the 12 KB pair table, for r = 5: 34 KB at
// 64 x 32, 52 KB at 64 x 64, 90 KB at 128 x 64) and, with the registers the
// tile's window takes, the blocks per CU (4 down to 1).
template <int TX, int TY, int R>
__global__ void __launch_bounds__(256) k_box_queue(const BoxPair *__restrict__ pairs, int npairs, unsigned ntiles,
                                                   unsigned *__restrict__ counter) {
    constexpr int E = 2 * R + 1, TH = TY + 2 * R, TW = TX + 2 * R;
    constexpr int NB = (TH * TW + 255) / 256;
    constexpr int NG = 256 / TX;   // row groups of the vertical pass
    static_assert(TX == 64 || TX == 128, "tile width");
    static_assert(TY % NG == 0, "strip rows must split over the row groups");
    __shared__ float tile[TH][TW + 1];
    __shared__ float hs[TH][TX + 1];
    __shared__ BoxPair sp[BQ_MAXP];   // the pair table: decoding an item reads no global memory
    __shared__ unsigned item;
    const int tx = threadIdx.x & (TX - 1), ty = threadIdx.x / TX;
    for (int i = threadIdx.x; i < npairs; i += 256) sp[i] = pairs[i];
    auto decode = [&](unsigned t) {
        int lo = 0, hi = npairs - 1;   // last pair whose first item is <= t
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (sp[mid].tile0 <= t) lo = mid;
            else hi = mid - 1;
        }
        const BoxPair bp = sp[lo];
        const unsigned u = t - bp.tile0;
        const int xb = (int)(u % bp.nx), sb = (int)((u / bp.nx) % bp.ny), d = (int)(u / ((unsigned)bp.nx * bp.ny));
        const size_t base = (size_t)d * plane(bp.W, bp.H);
        BoxItem q;
        q.src = bp.in + base;
        q.dst = bp.out + base;
        q.W = bp.W;
        q.H = bp.H;
        q.x0 = xb * TX;
        q.y0 = sb * TY;
        q.inner = q.x0 - R >= 0 && q.x0 + TX + R <= q.W && q.y0 - R >= 0 && q.y0 + TY + R <= q.H;
        return q;
    };
    float v[NB];
    auto load_tile = [&](const BoxItem &q) {
#pragma unroll
        for (int k = 0; k < NB; k++) {
            const int i = threadIdx.x + 256 * k;
            const int rr = i / TW, cc = i - rr * TW;
            float x = 0.0f;
            if (rr < TH) {
                const int yy = q.y0 - R + rr, xx = q.x0 - R + cc;
                x = q.inner ? q.src[(size_t)yy * q.W + xx] : q.src[(size_t)wrapn(yy, q.H) * q.W + wrapn(xx, q.W)];
            }
            v[k] = x;
        }
    };
    // Software pipeline: the next item's tile is loaded into registers
    // while the current one is summed, and the item after that is requested
    // from the counter one iteration ahead.
    unsigned nxt = 0;
    if (threadIdx.x == 0) {
        item = atomicAdd(counter, 1u);
        nxt = atomicAdd(counter, 1u);
    }
    __syncthreads();   // also: the pair table is complete
    unsigned t = item;
    __syncthreads();   // every wave has read item before the loop rewrites it
    BoxItem cur = {};
    if (t < ntiles) {
        cur = decode(t);
        load_tile(cur);
    }
    while (t < ntiles) {
#pragma unroll
        for (int k = 0; k < NB; k++) {
            const int i = threadIdx.x + 256 * k;
            const int rr = i / TW, cc = i - rr * TW;
            if (rr < TH) tile[rr][cc] = v[k];
        }
        if (threadIdx.x == 0) {
            item = nxt;
            nxt = atomicAdd(counter, 1u);
        }
        __syncthreads();   // the tile is staged and the next item published
        const unsigned tn = item;
        BoxItem nq = {};
        if (tn < ntiles) {
            nq = decode(tn);
            load_tile(nq);
        }
        // Horizontal sums: a thread takes a 16-column piece of one row, keeps
        // its 16 + 2R inputs in registers and adds each output's E taps in
        // column order (one LDS read per input instead of E per output).
        for (int task = threadIdx.x; task < TH * (TX / 16); task += 256) {
            const int rr = task / (TX / 16), c0 = (task % (TX / 16)) * 16;
            float w[16 + 2 * R];
#pragma unroll
            for (int i = 0; i < 16 + 2 * R; i++) w[i] = tile[rr][c0 + i];
#pragma unroll
            for (int j = 0; j < 16; j++) {
                float h = 0.0f;
#pragma unroll
                for (int dx = 0; dx < E; dx++) h += w[j + dx];
                hs[rr][c0 + j] = h;
            }
        }
        __syncthreads();
        // Vertical sums: a thread takes TY/NG consecutive output rows of one
        // column, its TY/NG + 2R horizontal sums in registers, taps in row
        // order.
        constexpr int RPT = TY / NG;
        const int x = cur.x0 + tx;
        if (x < cur.W) {
            float cv[RPT + 2 * R];
#pragma unroll
            for (int i = 0; i < RPT + 2 * R; i++) cv[i] = hs[ty * RPT + i][tx];
#pragma unroll
            for (int j = 0; j < RPT; j++) {
                const int y = cur.y0 + ty * RPT + j;
                if (y < cur.H) {
                    float acc = 0.0f;
#pragma unroll
                    for (int dy = 0; dy < E; dy++) acc += cv[j + dy];
                    stream_store(&cur.dst[(size_t)y * cur.W + x], div_count(acc, (uint32_t)(E * E)));
                }
            }
        }
        __syncthreads();   // tile, hs and item reads are done before they are rewritten
        t = tn;
        cur = nq;
    }
}


__global__ void k_zero_u32(unsigned *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

struct PairBufs {
    uint8_t *gray[2], *rank[2];
    float *logimg[2], *vol, *agg;
    void *ncc_ws;   // the cost kernel's per-pixel records of both views
};

struct Chunk {
    std::vector<int> items;              // indices into the run's item list
    std::vector<rts_batch_item> key;     // what the graph was captured for
    std::vector<PairBufs> bufs;          // carved from the batch's arena
    BoxPair *dpairs = nullptr;           // device table for the persistent box
    size_t dpairs_cap = 0;
    unsigned *counter = nullptr;
    unsigned ntiles = 0;
    int box_tx = 0, box_ty = 0;
    int cost_dch = 0;   // disparities per cost block of the captured graph
    hipGraphExec_t exec = nullptr;
};

size_t pair_bytes(const rts_batch_item &it) {
    const size_t P = (size_t)it.width * it.height;
    const size_t V = P * it.disp * sizeof(float);
    auto al = [](size_t n) { return (n + 255) & ~(size_t)255; };
    return 2 * al(V) + 2 * (al(P) + al(P) + al(P * 4)) + al(rts_cost_rank_log_ncc_ws_bytes(it.width, it.height));
}

bool same_items(const std::vector<rts_batch_item> &a, const rts_batch_item *b, const std::vector<int> &idx) {
    if (a.size() != idx.size()) return false;
    for (size_t i = 0; i < idx.size(); i++)
        if (memcmp(&a[i], &b[idx[i]], sizeof(rts_batch_item)) != 0) return false;
    return true;
}

}  // namespace

// A chunk of the pipeline path: its pairs and the graph of their launches.
struct PipeChunk {
    std::vector<int> items;
    std::vector<rts_batch_item> key;
    hipGraphExec_t exec = nullptr;
};

typedef std::tuple<int, int, int> Shape;   // W, H, D

struct rts_batch {
    rts_pipeline_desc base{};
    size_t budget = 0;
    // the pipeline path: resident pipelines by shape, and its chunks
    bool generic = false;
    std::map<Shape, rts_pipeline *> pipes;
    std::vector<std::unique_ptr<PipeChunk>> pchunks;
    // One arena, sized for the largest chunk, serves every chunk in turn:
    // chunks run one after another on the stream, so their graphs may all
    // point into it.
    uint8_t *arena = nullptr;
    size_t arena_bytes = 0;
    int use_graphs = 1;
    int tuned = 0, box_tx = 64, box_ty = 32, cost_dch = 16;
    float tune_ms[9] = {};
    std::vector<std::unique_ptr<Chunk>> chunks;
    rts_batch_stats stats{};
};

namespace {

void free_chunk(Chunk *c) {
    if (c->exec) (void)hipGraphExecDestroy(c->exec);
    (void)hipFree(c->dpairs);
    (void)hipFree(c->counter);
    c->exec = nullptr;
    c->dpairs = nullptr;
    c->dpairs_cap = 0;
    c->counter = nullptr;
}

// The strip height whose staged rows (TY + 2R) split into at most 256
// sixteen-column pieces, so the horizontal pass is one round of the block:
// 64 - 2R rounded down to a multiple of 4 (52 rows for the 11x11 box).
constexpr int box_balanced_rows(int r) { return ((64 - 2 * r) / 4) * 4; }

template <int TX, int R>
void launch_box_w(const Chunk &c, int ty, int nb, hipStream_t s) {
    constexpr int TB = box_balanced_rows(R);
    if (ty == 32)
        k_box_queue<TX, 32, R><<<nb, 256, 0, s>>>(c.dpairs, (int)c.bufs.size(), c.ntiles, c.counter);
    else if (ty == TB)
        k_box_queue<TX, TB, R><<<nb, 256, 0, s>>>(c.dpairs, (int)c.bufs.size(), c.ntiles, c.counter);
    else
        k_box_queue<TX, 64, R><<<nb, 256, 0, s>>>(c.dpairs, (int)c.bufs.size(), c.ntiles, c.counter);
}

template <int R>
void launch_box_r(const Chunk &c, int tx, int ty, int nb, hipStream_t s) {
    if (tx == 128)
        launch_box_w<128, R>(c, ty, nb, s);
    else
        launch_box_w<64, R>(c, ty, nb, s);
}

// The box over the chunk's work table, built for tiles of tx x ty.
int launch_box(const Chunk &c, int r, hipStream_t s) {
    const int tx = c.box_tx, ty = c.box_ty;
    const int nb = std::max(1, std::min<int>((int)std::min<unsigned>(c.ntiles, 1u << 30), 256 * 8));
    k_zero_u32<<<1, 64, 0, s>>>(c.counter, 1);
    RTS_LAUNCHED();
    switch (r) {
        case 0: launch_box_r<0>(c, tx, ty, nb, s); break;
        case 1: launch_box_r<1>(c, tx, ty, nb, s); break;
        case 2: launch_box_r<2>(c, tx, ty, nb, s); break;
        case 3: launch_box_r<3>(c, tx, ty, nb, s); break;
        case 4: launch_box_r<4>(c, tx, ty, nb, s); break;
        case 5: launch_box_r<5>(c, tx, ty, nb, s); break;
        case 6: launch_box_r<6>(c, tx, ty, nb, s); break;
        case 7: launch_box_r<7>(c, tx, ty, nb, s); break;
        default: launch_box_r<BQ_RMAX>(c, tx, ty, nb, s); break;
    }
    RTS_LAUNCHED();
    return RTS_OK;
}

// Sets the chunk's box work table for tiles of tx columns x ty rows (device
// upload).
int build_box_table(Chunk &c, const rts_batch_item *items, int tx, int ty, hipStream_t s) {
    std::vector<BoxPair> t(c.items.size());
    unsigned n = 0;
    for (size_t i = 0; i < c.items.size(); i++) {
        const rts_batch_item &it = items[c.items[i]];
        BoxPair &b = t[i];
        b.in = c.bufs[i].vol;
        b.out = c.bufs[i].agg;
        b.W = it.width;
        b.H = it.height;
        b.D = it.disp;
        b.nx = div_up(it.width, tx);
        b.ny = div_up(it.height, ty);
        b.tile0 = n;
        const unsigned long add = (unsigned long)b.nx * b.ny * b.D;
        RTS_CHECK(n + add < (1ul << 31), "too many box work items in one chunk");
        n += (unsigned)add;
    }
    if (c.dpairs_cap < t.size()) {
        (void)hipFree(c.dpairs);
        c.dpairs = nullptr;
        c.dpairs_cap = 0;
        RTS_HIP(hipMalloc((void **)&c.dpairs, sizeof(BoxPair) * t.size()));
        c.dpairs_cap = t.size();
    }
    if (!c.counter) RTS_HIP(hipMalloc((void **)&c.counter, 256));
    RTS_HIP(hipMemcpyAsync(c.dpairs, t.data(), sizeof(BoxPair) * t.size(), hipMemcpyHostToDevice, s));
    RTS_HIP(hipStreamSynchronize(s));
    c.ntiles = n;
    c.box_tx = tx;
    c.box_ty = ty;
    return RTS_OK;
}

int cost_pair(const rts_batch *b, const rts_batch_item &it, const PairBufs &pb, int dch, hipStream_t s) {
    const rts_pipeline_desc &d = b->base;
    return rts_cost_rank_log_ncc_ws(pb.gray[0], pb.gray[1], pb.rank[0], pb.rank[1], pb.logimg[0], pb.logimg[1], pb.vol,
                                    it.width, it.height, it.disp, d.ncc_hr, d.ncc_wr, d.lambda1, d.lambda2, d.lambda3,
                                    dch, pb.ncc_ws, rts_cost_rank_log_ncc_ws_bytes(it.width, it.height), s);
}

int prep_pair(const rts_batch *b, const rts_batch_item &it, const PairBufs &pb, hipStream_t s) {
    static const float kLoG5[25] = {0.0239f, 0.046f,   0.0499f, 0.046f,   0.0239f,  0.046f,  0.0061f, -0.0923f, 0.0061f,
                                    0.046f,  0.0499f,  -0.0923f, -0.3182f, -0.0923f, 0.0499f, 0.046f,  0.0061f, -0.0923f,
                                    0.0061f, 0.046f,   0.0239f, 0.046f,   0.0499f,  0.046f,  0.0239f};
    const uint8_t *rgb[2] = {it.left_rgb, it.right_rgb};
    for (int v = 0; v < 2; v++) {
        if (int e = rts_rgb_to_gray(rgb[v], pb.gray[v], it.width, it.height, s)) return e;
        if (int e = rts_rank(pb.gray[v], pb.rank[v], it.width, it.height, b->base.rank_m, s)) return e;
        if (int e = rts_convolve(pb.gray[v], pb.logimg[v], kLoG5, it.width, it.height, 2, 2, s)) return e;
    }
    return RTS_OK;
}

// The chunk's whole launch sequence on stream s.
int enqueue_chunk(rts_batch *b, Chunk &c, const rts_batch_item *items, hipStream_t s) {
    for (size_t i = 0; i < c.items.size(); i++) {
        const rts_batch_item &it = items[c.items[i]];
        if (int e = prep_pair(b, it, c.bufs[i], s)) return e;
        if (int e = cost_pair(b, it, c.bufs[i], b->cost_dch, s)) return e;
    }
    if (int e = launch_box(c, b->base.box_r, s)) return e;
    for (size_t i = 0; i < c.items.size(); i++) {
        const rts_batch_item &it = items[c.items[i]];
        if (int e = rts_wta(c.bufs[i].agg, it.dl, it.dr, it.mask, it.width, it.height, it.disp, 0, b->base.lr_tol, s))
            return e;
    }
    return RTS_OK;
}

// Carves the chunk's pair buffers out of the arena (c.items set).
void carve(rts_batch *b, Chunk &c, const rts_batch_item *items) {
    auto al = [](size_t n) { return (n + 255) & ~(size_t)255; };
    uint8_t *p = b->arena;
    c.bufs.resize(c.items.size());
    for (size_t i = 0; i < c.items.size(); i++) {
        const rts_batch_item &it = items[c.items[i]];
        const size_t P = (size_t)it.width * it.height, V = P * it.disp * sizeof(float);
        PairBufs &b = c.bufs[i];
        b.vol = (float *)p;
        p += al(V);
        b.agg = (float *)p;
        p += al(V);
        for (int v = 0; v < 2; v++) {
            b.gray[v] = p;
            p += al(P);
            b.rank[v] = p;
            p += al(P);
            b.logimg[v] = (float *)p;
            p += al(P * 4);
        }
        b.ncc_ws = p;
        p += al(rts_cost_rank_log_ncc_ws_bytes(it.width, it.height));
    }
}

// Times the candidates of each tuned parameter over every pair of the first
// chunk (the best of the timed runs after a warm-up run): the cost kernel's
// disparities per block (8, 16, 32), one launch per pair, and the box tile:
// 64 or 128 columns by a strip of 32 rows, the balanced height of
// box_balanced_rows or 64 rows, one persistent launch.  The tile sets the
// LDS footprint (for r = 5 from 34 KB at 64 x 32 to 90 KB at 128 x 64) and so
// the blocks per CU, and the number of work items.  The chunk's mix matters:
// on the largest pair the cost candidates nearly tie, while larger tiles
// leave the small pairs' launches with few blocks.  Candidates are tried
// from the smallest footprint up; a larger one replaces a smaller one only
// when it is 2% faster.
int autotune(rts_batch *b, Chunk &c, const rts_batch_item *items, hipStream_t s) {
    hipEvent_t e0, e1;
    RTS_HIP(hipEventCreate(&e0));
    RTS_HIP(hipEventCreate(&e1));
    int st = RTS_OK;
    for (size_t i = 0; i < c.items.size() && st == RTS_OK; i++) st = prep_pair(b, items[c.items[i]], c.bufs[i], s);
    const int dchs[3] = {8, 16, 32};
    const int tb = box_balanced_rows(b->base.box_r);
    const int box_tx[6] = {64, 64, 64, 128, 128, 128}, box_ty[6] = {32, tb, 64, 32, tb, 64};
    constexpr int REPS = 3;            // the first run warms up
    constexpr float SWITCH = 0.98f;    // a larger candidate must be 2% faster
    // times run(k) as the best of REPS - 1 runs after a warm-up
    auto time_it = [&](auto run) -> float {
        float ms = 1e30f;
        for (int rep = 0; rep < REPS && st == RTS_OK; rep++) {
            float t = 0.0f;
            st = hipEventRecord(e0, s) == hipSuccess ? RTS_OK : fail(RTS_ERR_HIP, "event");
            if (st == RTS_OK) st = run();
            if (st == RTS_OK) st = hipEventRecord(e1, s) == hipSuccess ? RTS_OK : fail(RTS_ERR_HIP, "event");
            if (st == RTS_OK && hipEventSynchronize(e1) == hipSuccess) (void)hipEventElapsedTime(&t, e0, e1);
            if (rep > 0 && t < ms) ms = t;
        }
        return ms;
    };
    float best = 1e30f;
    for (int k = 0; k < 3 && st == RTS_OK; k++) {
        const float ms = time_it([&]() {
            for (size_t i = 0; i < c.items.size(); i++)
                if (int e = cost_pair(b, items[c.items[i]], c.bufs[i], dchs[k], s)) return e;
            return (int)RTS_OK;
        });
        b->tune_ms[k] = ms;
        if (k == 0 || ms < best * SWITCH) {
            best = ms;
            b->cost_dch = dchs[k];
        }
    }
    best = 1e30f;
    for (int k = 0; k < 6 && st == RTS_OK; k++) {
        st = build_box_table(c, items, box_tx[k], box_ty[k], s);
        if (st != RTS_OK) break;
        const float ms = time_it([&]() { return launch_box(c, b->base.box_r, s); });
        b->tune_ms[3 + k] = ms;
        if (k == 0 || ms < best * SWITCH) {
            best = ms;
            b->box_tx = box_tx[k];
            b->box_ty = box_ty[k];
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->tuned = st == RTS_OK;
    return st;
}

rts_pipeline_desc shape_desc(const rts_batch *b, const Shape &k) {
    rts_pipeline_desc d = b->base;
    d.width = std::get<0>(k);
    d.height = std::get<1>(k);
    d.disp = std::get<2>(k);
    return d;
}

size_t shape_bytes(const rts_batch *b, const Shape &k) {
    const rts_pipeline_desc d = shape_desc(b, k);
    return rts_pipeline_bytes(&d);
}

void drop_pchunk(PipeChunk *c) {
    if (c->exec) (void)hipGraphExecDestroy(c->exec);
    c->exec = nullptr;
}

// The pipeline path of rts_batch_run.
int run_generic(rts_batch *b, const rts_batch_item *items, int n, hipStream_t s) {
    // chunks: consecutive pairs while the pipelines of their distinct shapes
    // fit the budget
    std::vector<std::vector<int>> plan;
    std::vector<std::set<Shape>> plan_shapes;
    size_t used = 0;
    for (int i = 0; i < n; i++) {
        const Shape k{items[i].width, items[i].height, items[i].disp};
        const size_t need = shape_bytes(b, k);
        RTS_CHECK(need <= b->budget, "pair %d needs %zu bytes, over the budget of %zu", i, need, b->budget);
        const bool have = !plan.empty() && plan_shapes.back().count(k);
        if (plan.empty() || (!have && used + need > b->budget)) {
            plan.emplace_back();
            plan_shapes.emplace_back();
            used = 0;
        }
        plan.back().push_back(i);
        if (!have) {
            plan_shapes.back().insert(k);
            used += need;
        }
    }
    while (b->pchunks.size() < plan.size()) b->pchunks.emplace_back(new PipeChunk());
    b->stats.pairs = n;
    b->stats.chunks = (int)plan.size();
    for (size_t c = 0; c < plan.size(); c++) {
        PipeChunk &ch = *b->pchunks[c];
        // make room: pipelines of other shapes leave while the resident set
        // with this chunk's new shapes would exceed the budget (every graph
        // that used them goes with them)
        size_t resident = 0, add = 0;
        for (auto &kv : b->pipes) resident += shape_bytes(b, kv.first);
        for (const Shape &k : plan_shapes[c])
            if (!b->pipes.count(k)) add += shape_bytes(b, k);
        if (resident + add > b->budget) {
            RTS_HIP(hipStreamSynchronize(s));
            for (auto it = b->pipes.begin(); it != b->pipes.end() && resident + add > b->budget;) {
                if (plan_shapes[c].count(it->first)) {
                    ++it;
                    continue;
                }
                resident -= shape_bytes(b, it->first);
                rts_pipeline_destroy(it->second);
                it = b->pipes.erase(it);
                for (auto &pc : b->pchunks) {
                    drop_pchunk(pc.get());
                    pc->key.clear();
                }
            }
        }
        for (const Shape &k : plan_shapes[c])
            if (!b->pipes.count(k)) {
                const rts_pipeline_desc d = shape_desc(b, k);
                rts_pipeline *p = nullptr;
                if (int e = rts_pipeline_create(&d, &p)) return e;
                b->pipes[k] = p;
            }
        auto enqueue = [&]() -> int {
            for (int i : plan[c]) {
                const rts_batch_item &it = items[i];
                rts_pipeline *p = b->pipes[Shape{it.width, it.height, it.disp}];
                if (int e = pipeline_enqueue(p, it.left_rgb, it.right_rgb, it.dl, it.dr, it.mask, s)) return e;
            }
            return (int)RTS_OK;
        };
        if (!b->use_graphs) {
            if (int e = enqueue()) return e;
            continue;
        }
        if (ch.exec && same_items(ch.key, items, plan[c])) {
            RTS_HIP(hipGraphLaunch(ch.exec, s));
            b->stats.graph_replays++;
            continue;
        }
        drop_pchunk(&ch);
        ch.items = plan[c];
        ch.key.clear();
        for (int i : ch.items) ch.key.push_back(items[i]);
        hipGraph_t g = nullptr;
        RTS_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        const int e = enqueue();
        const hipError_t ce = hipStreamEndCapture(s, &g);
        if (e != RTS_OK) {
            if (g) (void)hipGraphDestroy(g);
            ch.key.clear();
            return e;
        }
        RTS_HIP(ce);
        const hipError_t ie = hipGraphInstantiate(&ch.exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ie != hipSuccess) {
            ch.exec = nullptr;
            ch.key.clear();
            RTS_HIP(ie);
        }
        b->stats.graphs_built++;
        RTS_HIP(hipGraphLaunch(ch.exec, s));
    }
    size_t resident = 0;
    for (auto &kv : b->pipes) resident += shape_bytes(b, kv.first);
    b->stats.arena_bytes = resident;
    return RTS_OK;
}

}  // namespace

extern "C" {

rts_status rts_batch_create(const rts_pipeline_desc *base, size_t budget_bytes, int use_graphs, rts_batch **out) {
    RTS_CHECK(base && out, "null argument");
    const bool special = base->cost1 == RTS_COST_RANK && base->cost2 == RTS_COST_LOG && base->cost3 == RTS_COST_NCC &&
                         base->aggr == RTS_AGGR_BOX;
    if (special) {
        RTS_CHECK(base->box_r >= 0 && base->box_r <= BQ_RMAX, "box radius %d must be in [0, %d]", base->box_r,
                  BQ_RMAX);
        RTS_CHECK(base->lambda1 > 0 && base->lambda2 > 0 && base->lambda3 > 0, "combine weights must be positive");
    } else {
        // the configuration must make a pipeline (checked at a size every
        // size check accepts; each pair's size is checked when it runs)
        rts_pipeline_desc d = *base;
        d.width = 64;
        d.height = 64;
        d.disp = 16;
        if (int e = pipeline_check(d)) return e;
    }
    RTS_CHECK(budget_bytes > 0, "empty budget");
    rts_batch *b = new (std::nothrow) rts_batch();
    RTS_CHECK(b != nullptr, "out of host memory");
    b->base = *base;
    b->generic = !special;
    b->budget = budget_bytes;
    b->use_graphs = use_graphs != 0;
    *out = b;
    return RTS_OK;
}

rts_status rts_batch_destroy(rts_batch *b) {
    if (!b) return RTS_OK;
    for (auto &c : b->chunks) free_chunk(c.get());
    for (auto &c : b->pchunks) drop_pchunk(c.get());
    for (auto &kv : b->pipes) rts_pipeline_destroy(kv.second);
    (void)hipFree(b->arena);
    delete b;
    return RTS_OK;
}

rts_status rts_batch_run(rts_batch *b, const rts_batch_item *items, int n, void *stream) {
    RTS_CHECK(b && (items || n == 0), "null argument");
    RTS_CHECK(n >= 0, "negative pair count");
    hipStream_t s = S(stream);
    if (b->generic) {
        for (int i = 0; i < n; i++) {
            const rts_batch_item &it = items[i];
            if (int e = check_dims(it.width, it.height, it.disp)) return e;
            RTS_CHECK(it.left_rgb && it.right_rgb && it.dl && (!b->base.lr_check || (it.dr && it.mask)),
                      "pair %d: null buffer", i);
        }
        RTS_CHECK(!b->use_graphs || stream != nullptr, "graph capture needs a stream other than the null stream");
        return run_generic(b, items, n, s);
    }
    for (int i = 0; i < n; i++) {
        const rts_batch_item &it = items[i];
        if (int e = check_dims(it.width, it.height, it.disp)) return e;
        RTS_CHECK(it.left_rgb && it.right_rgb && it.dl && it.dr && it.mask, "pair %d: null buffer", i);
        RTS_CHECK(2 * b->base.ncc_hr + 1 <= it.height && 2 * b->base.ncc_wr + 1 <= it.width &&
                      b->base.box_r < it.width && b->base.box_r < it.height,
                  "pair %d: windows larger than the image", i);
        RTS_CHECK(pair_bytes(it) <= b->budget, "pair %d needs %zu bytes, over the budget of %zu", i, pair_bytes(it),
                  b->budget);
    }
    // Chunks: consecutive pairs while their footprint fits the budget.
    std::vector<std::vector<int>> plan;
    size_t used = 0, arena_need = 0;
    for (int i = 0; i < n; i++) {
        const size_t pb = pair_bytes(items[i]);
        if (plan.empty() || used + pb > b->budget || plan.back().size() >= (size_t)BQ_MAXP) {
            plan.emplace_back();
            used = 0;
        }
        plan.back().push_back(i);
        used += pb;
        arena_need = std::max(arena_need, used);
    }
    RTS_CHECK(!b->use_graphs || stream != nullptr, "graph capture needs a stream other than the null stream");
    if (b->arena_bytes < arena_need) {   // a new arena invalidates every captured graph
        for (auto &c : b->chunks) free_chunk(c.get());
        b->chunks.clear();
        RTS_HIP(hipStreamSynchronize(s));
        (void)hipFree(b->arena);
        b->arena = nullptr;
        b->arena_bytes = 0;
        if (hipMalloc((void **)&b->arena, arena_need) != hipSuccess) {
            (void)hipGetLastError();
            return fail(RTS_ERR_NOMEM, "batch: arena of %zu bytes", arena_need);
        }
        b->arena_bytes = arena_need;
    }
    while (b->chunks.size() < plan.size()) b->chunks.emplace_back(new Chunk());
    b->stats.pairs = n;
    b->stats.chunks = (int)plan.size();
    b->stats.arena_bytes = b->arena_bytes;
    for (size_t k = 0; k < plan.size(); k++) {
        Chunk &c = *b->chunks[k];
        const bool reuse = c.exec && c.box_tx == b->box_tx && c.box_ty == b->box_ty && c.cost_dch == b->cost_dch &&
                           same_items(c.key, items, plan[k]);
        if (!reuse) {
            if (c.exec) (void)hipGraphExecDestroy(c.exec);
            c.exec = nullptr;
            c.items = plan[k];
            c.key.clear();
            for (int i : c.items) c.key.push_back(items[i]);
            carve(b, c, items);
            if (!b->tuned)
                if (int e = autotune(b, c, items, s)) return e;
            if (int e = build_box_table(c, items, b->box_tx, b->box_ty, s)) return e;
            c.cost_dch = b->cost_dch;
            if (b->use_graphs) {
                hipGraph_t g = nullptr;
                RTS_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                const int e = enqueue_chunk(b, c, items, s);
                const hipError_t ce = hipStreamEndCapture(s, &g);
                if (e != RTS_OK) {
                    if (g) (void)hipGraphDestroy(g);
                    return e;
                }
                RTS_HIP(ce);
                const hipError_t ie = hipGraphInstantiate(&c.exec, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                RTS_HIP(ie);
                b->stats.graphs_built++;
            }
        } else {
            b->stats.graph_replays++;
        }
        if (b->use_graphs) {
            RTS_HIP(hipGraphLaunch(c.exec, s));
        } else {
            if (int e = enqueue_chunk(b, c, items, s)) return e;
        }
        b->stats.box_items += c.ntiles;
    }
    b->stats.box_rows = b->box_ty;
    b->stats.box_cols = b->box_tx;
    b->stats.cost_disparities = b->cost_dch;
    for (int i = 0; i < 9; i++) b->stats.tune_ms[i] = b->tune_ms[i];
    return RTS_OK;
}

rts_status rts_batch_set_tiles(rts_batch *b, int box_cols, int box_rows, int cost_disparities) {
    RTS_CHECK(b != nullptr, "null batch");
    if (box_cols == 0 && box_rows == 0 && cost_disparities == 0) {
        b->tuned = 0;   // time the candidates again on the next run's first chunk
        return RTS_OK;
    }
    const int tb = box_balanced_rows(b->base.box_r);
    RTS_CHECK(box_cols == 64 || box_cols == 128, "box tile width %d must be 64 or 128", box_cols);
    RTS_CHECK(box_rows == 32 || box_rows == tb || box_rows == 64, "box strip of %d rows must be 32, %d or 64", box_rows,
              tb);
    RTS_CHECK(cost_disparities == 8 || cost_disparities == 16 || cost_disparities == 32,
              "disparities per cost block %d must be 8, 16 or 32", cost_disparities);
    b->box_tx = box_cols;
    b->box_ty = box_rows;
    b->cost_dch = cost_disparities;
    b->tuned = 1;
    return RTS_OK;
}

rts_status rts_batch_get_stats(rts_batch *b, rts_batch_stats *st) {
    RTS_CHECK(b && st, "null argument");
    *st = b->stats;
    return RTS_OK;
}

size_t rts_batch_pair_bytes(int W, int H, int D) {
    rts_batch_item it{};
    it.width = W;
    it.height = H;
    it.disp = D;
    return pair_bytes(it);
}

}  // extern "C"
