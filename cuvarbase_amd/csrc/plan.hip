// Batch planning and execution (C ABI).
//
// A plan fixes, for every curve of a batch, the grid size of its three NFFTs
// (n = n1 * 2^p2 >= sigma*nf from the size catalogue), the four-step split,
// the Gaussian width b and the workspace layout.  Curves with the same split
// form a size bucket; each bucket runs as two launches (pass 1 and pass 2)
// after one prep launch for the whole batch.  This replaces the reference's
// per-curve allocation, two cuFFT plans per curve and nine launches per curve
// (cuvarbase/lombscargle.py:129-181, :579-689; cuvarbase/cunfft.py:82-92,
// :163-308).
//
// A plan is a host object only: it owns no device memory.  Its per-curve
// descriptors are copied (hipMemcpyAsync, on the execution stream) into the
// first bytes of the caller's workspace at every execute, so plans can be
// created and destroyed while other batches run without any device
// allocation or the implicit synchronisation of hipFree.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <atomic>
#include <map>
#include <mutex>
#include <tuple>

#include "gls_internal.hpp"

struct gls_plan {
    uint64_t id = 0;       // unique per plan (graphs remember the plan they came from)
    int kind = 0;  // 0 lomb, 1 nfft
    int64_t pts_cap = 0;   // points the per-point workspace regions hold
    gls::PlanParams pp{};
    double sigma = 0.0;
    int ncurves = 0;
    std::vector<gls::CurveGeom> geom;
    struct Bucket {
        int n1 = 0, p2 = 0, c1 = 0, cr = 0;   // size and tiles (columns, rows)
        int64_t first = 0;  // offset (curves) into the bucket curve list
        int count = 0;
        int64_t bkt0 = 0;   // bucket offsets of its first curve (the others follow)
        bool fix = false;   // fixed-point spreading in pass 1
        bool nts = false;   // pass 1 stores the grids non-temporally
        // the next entry is the same size bucket in the other spreading mode:
        // one pass-2 launch after both entries' pass 1 serves the two
        bool pass2_with_next = false;
    };
    std::vector<Bucket> buckets;   // one entry per chunk of a size bucket and spread mode
    int spread_mode = GLS_SPREAD_FLOAT;  // gls_params.spread
    std::vector<int8_t> curve_fixed;     // per curve: pass 1 spreads it in fixed point
    std::vector<int> bucket_lane;  // 0: the caller's stream, j > 0: auxiliary stream j-1
    bool fork_buckets = false;     // run the lanes concurrently (gls_plan_set_fork)
    std::vector<hipEvent_t> fork_join;  // fork event + one join event per aux stream
    std::vector<int32_t> bcurves;  // curve indices of the entries, in order
    int64_t chunk_bytes = 0;       // > 0: grid region reused by every entry
    int ntiles = 0;         // direct method
    int rf = 1;
    // host image of the descriptors uploaded into workspace[0 .. meta.size())
    std::vector<char> meta;
    size_t off_geom = 0, off_bgeom = 0, off_tiles = 0, off_prep = 0;
    std::vector<int32_t> prep_list;   // curves prepared by 256-thread blocks, then wide ones
    int nprep_small = 0;
    size_t off_state = 0, off_nd0 = 0, off_nd1 = 0, off_nd2 = 0,
           off_w = 0, off_yw = 0, off_dt = 0, off_bkt = 0,
           off_grid = 0;
    // per wide curve bucketing records (NFFT method; see WideAux) and chunk
    // partials of their reductions (WidePart)
    size_t off_wide = 0, off_wpart = 0;
    int nwide_aux = 0;
    int wide_max_obs = 0;
    size_t ws_bytes = 0;
    int64_t total_out = 0;
    // optional per-stage timing (events recorded between launches)
    struct Stage {
        int kind;      // 0 prep, 1 pass 1, 2 pass 2, 3 direct sums
        int grid;      // grid size n of the bucket (0 for prep / direct)
        double bytes;  // algorithmic bytes moved by the stage
        int bucket;    // index into buckets (passes 1 and 2), else -1
    };
    std::vector<Stage> stages;
    bool timing = false;
    std::vector<hipEvent_t> events;
};

namespace gls {
namespace {

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

std::atomic<uint64_t> g_plan_ids{1};

struct Layout {
    size_t cur = 0;
    size_t take(size_t bytes) {
        const size_t o = cur;
        cur = align_up(cur + bytes, 256);
        return o;
    }
};

int radix_stages(int L) {
    int st = 0, rem = L;
    const int radices[] = {16, 12, 9, 8, 4, 2, 5, 3};   // the order of next_radix
    while (rem > 1) {
        bool found = false;
        for (int r : radices)
            if (rem % r == 0) {
                rem /= r;
                ++st;
                found = true;
                break;
            }
        if (!found) return 99;
    }
    return st;
}

double gauss_b(double sigma, int m) {
    // b = 2 sigma m / ((2 sigma - 1) pi)  (cuvarbase/cunfft.py:56-57); for the
    // centred band this balances the alias term exp(-pi^2 b (1 - 1/sigma))
    // against the deconvolution-amplified truncation exp(-m^2/b + pi^2 b/(4 sigma^2))
    return 2.0 * sigma * (double)m / ((2.0 * sigma - 1.0) * M_PI);
}

// work-item cap of one NFFT-pass launch (gls_launch_items_set)
std::atomic<int64_t> g_max_launch_items{kMaxLaunchItems};

// ---- tiles and measured costs of the size buckets ------------------------------
std::mutex g_tile_mu;
std::map<std::pair<int, int>, BucketTiles> g_tile_override;   // gls_tile_set
std::map<std::pair<int, int>, double> g_cost_override;        // gls_grid_cost_set

constexpr int kNumRowSizes = kMaxLog2Row - kMinLog2Row + 1;

int size_index(int n1) {
    for (int i = 0; i < kNumPass1Sizes; ++i)
        if (kPass1Sizes[i] == n1) return i;
    return -1;
}

// model cost per cell of an unmeasured bucket: one pass over the cells plus
// 0.12 per LDS radix stage of the two FFTs, in the table's units
double model_cost(int n1, int p2) {
    return GLS_MODEL_PS_PER_UNIT * (1.0 + 0.12 * (radix_stages(n1) + radix_stages(1 << p2)));
}

// The measured table compiled in (tiles_gen.hpp) as arrays over (p2,
// catalogue position), built on first use: each bucket's tiles (default
// rules where the table has no row) and cost per cell (the model where it
// has none), so that planning does no table scans.
struct TableIndex {
    BucketTiles tiles[kNumRowSizes][kNumPass1Sizes];
    double ps[kNumRowSizes][kNumPass1Sizes];
    bool measured[kNumRowSizes][kNumPass1Sizes];
    TableIndex() {
        for (int p2 = kMinLog2Row; p2 <= kMaxLog2Row; ++p2)
            for (int i = 0; i < kNumPass1Sizes; ++i) {
                const int n1 = kPass1Sizes[i];
                tiles[p2 - kMinLog2Row][i] = BucketTiles{default_pass1_cols(n1),
                                                         default_pass2_rows(p2)};
                ps[p2 - kMinLog2Row][i] = model_cost(n1, p2);
                measured[p2 - kMinLog2Row][i] = false;
            }
#define GLS_TILE_ROW(N, P, C, R, PS) set(N, P, C, R, PS);
        GLS_TILE_TABLE(GLS_TILE_ROW)
#undef GLS_TILE_ROW
    }
    void set(int n1, int p2, int c1, int cr, double cost) {
        const int i = size_index(n1);
        if (i < 0 || p2 < kMinLog2Row || p2 > kMaxLog2Row) return;
        tiles[p2 - kMinLog2Row][i] = BucketTiles{c1, cr};
        ps[p2 - kMinLog2Row][i] = cost;
        measured[p2 - kMinLog2Row][i] = true;
    }
};
const TableIndex& table_index() {
    static const TableIndex index;   // thread-safe initialisation on first use
    return index;
}

BucketTiles table_tiles(int n1, int p2) {
    const int i = size_index(n1);
    if (i < 0 || p2 < kMinLog2Row || p2 > kMaxLog2Row)
        return BucketTiles{default_pass1_cols(n1), default_pass2_rows(p2)};
    return table_index().tiles[p2 - kMinLog2Row][i];
}

// cost per cell of every bucket, looked up once per plan: the caller's
// override (gls_grid_cost_set), else the measured table, else the model
struct GridCosts {
    double ps[kNumRowSizes][kNumPass1Sizes];
    GridCosts() {
        std::memcpy(ps, table_index().ps, sizeof(ps));
        std::lock_guard<std::mutex> lk(g_tile_mu);
        for (const auto& kv : g_cost_override) {
            const int i = size_index(kv.first.first);
            if (i >= 0) ps[kv.first.second - kMinLog2Row][i] = kv.second;
        }
    }
};

// the tiles plans made now use for bucket (n1, p2)
BucketTiles current_tiles(int n1, int p2) {
    BucketTiles t = table_tiles(n1, p2);
    std::lock_guard<std::mutex> lk(g_tile_mu);
    auto it = g_tile_override.find({n1, p2});
    if (it != g_tile_override.end()) {
        if (it->second.c1) t.c1 = it->second.c1;
        if (it->second.cr) t.cr = it->second.cr;
    }
    return t;
}

// the tiles of one plan's buckets, looked up once per bucket
struct TileCache {
    std::map<std::pair<int, int>, BucketTiles> seen;
    BucketTiles get(int n1, int p2) {
        auto it = seen.find({n1, p2});
        if (it != seen.end()) return it->second;
        return seen[{n1, p2}] = current_tiles(n1, p2);
    }
};

// cheapest grid n = n1 * 2^p2 >= need from the size catalogue, by cost per
// cell times cells (a point's 2m + 1 taps and a pass-1 tile must fit in one
// row: 2m + c1 <= n2)
bool choose_grid(int64_t need, int m, TileCache& tiles, const GridCosts& costs, int* n1_out,
                 int* p2_out) {
    double best = 0.0;
    bool found = false;
    if (need < kMinGrid) need = kMinGrid;
    for (int p2 = kMinLog2Row; p2 <= kMaxLog2Row; ++p2) {
        const int64_t n2 = 1ll << p2;
        for (int i = 0; i < kNumPass1Sizes; ++i) {
            const int n1 = kPass1Sizes[i];
            const int64_t n = (int64_t)n1 * n2;
            if (n < need || n > kMaxGrid) continue;
            if (2 * m + 16 > n2 && 2 * m + tiles.get(n1, p2).c1 > n2) continue;
            const double cost = (double)n * costs.ps[p2 - kMinLog2Row][i];
            if (!found || cost < best) {
                best = cost;
                *n1_out = n1;
                *p2_out = p2;
                found = true;
            }
        }
    }
    return found;
}

// planning of the NFFT grids of a batch; ngrids = 3 (lomb) or 1 (nfft).
// Curves with the same (n1, p2) form a size bucket; with chunking, a bucket is
// cut into entries of at most chunk_bytes of grids, every entry's grids start
// at offset 0 of one shared region (entries run one after the other on the
// stream), and the region is only as large as the largest entry.
int plan_nfft_geometry(gls_plan* P, int ngrids, Layout& L, int64_t total_pts,
                       std::vector<int32_t>& bcurves) {
    const int m = P->pp.m;
    TileCache tiles;
    const GridCosts costs;
    // (n1, p2, fixed-point spreading) -> curves
    std::map<std::tuple<int, int, int>, std::vector<int32_t>> by_size;
    P->curve_fixed.assign(P->ncurves, 0);
    int64_t bkt_ints = 0;
    for (int i = 0; i < P->ncurves; ++i) {
        CurveGeom& G = P->geom[i];
        const int64_t need = (int64_t)std::ceil(P->sigma * (double)G.nf);
        int n1 = 0, p2 = 0;
        if (!choose_grid(need, m, tiles, costs, &n1, &p2))
            return fail(GLS_E_UNSUPPORTED,
                        "no grid for curve " + std::to_string(i) + " (nf=" +
                            std::to_string(G.nf) + ", sigma*nf must be <= 2^24)");
        const int c1 = tiles.get(n1, p2).c1;
        const int n2 = 1 << p2;
        const int64_t n = (int64_t)n1 * n2;
        G.n = (int32_t)n;
        G.n1 = n1;
        G.p2 = p2;
        G.c1 = c1;
        G.kc = G.nf / 2;
        G.b = gauss_b((double)n / (double)G.nf, m);
        G.g_binv = (float)(1.4426950408889634 / G.b);
        G.g_norm = (float)(1.0 / std::sqrt(M_PI * G.b));
        G.g_bpi2 = (float)(G.b * M_PI * M_PI);
        G.ninv = 1.0f / (float)n;
        // GLS_SPREAD_AUTO: fixed point for curves of at least one point per
        // kFixedCellsPerPoint grid cells, where its integer adds outrun the
        // FP32 atomics of dense tiles, on columns of kFixedMinColumn cells or
        // more
        const int fix = P->spread_mode == GLS_SPREAD_FIXED ||
                        (P->spread_mode == GLS_SPREAD_AUTO && n1 >= kFixedMinColumn &&
                         (int64_t)G.n_obs * kFixedCellsPerPoint >= n);
        P->curve_fixed[i] = (int8_t)fix;
        by_size[{n1, p2, fix}].push_back(i);
    }
    const int64_t chunk = P->chunk_bytes;
    int64_t grid_total = 0, grid_region = 0;
    for (auto& kv : by_size) {
        gls_plan::Bucket B;
        B.n1 = std::get<0>(kv.first);
        B.p2 = std::get<1>(kv.first);
        B.fix = std::get<2>(kv.first) != 0;
        const BucketTiles bt = tiles.get(B.n1, B.p2);
        B.c1 = bt.c1;
        B.cr = bt.cr;
        B.first = (int64_t)bcurves.size();
        B.count = 0;
        B.bkt0 = bkt_ints;
        int64_t acc = 0;   // grid elements of the open entry
        for (int32_t ci : kv.second) {
            CurveGeom& G = P->geom[ci];
            const int64_t cells = (int64_t)ngrids * grid_cells(G.n1, G.p2);
            if (chunk > 0 && B.count > 0 && (acc + cells) * (int64_t)sizeof(float2) > chunk) {
                P->buckets.push_back(B);
                B.first = (int64_t)bcurves.size();
                B.count = 0;
                B.bkt0 = bkt_ints;
                acc = 0;
            }
            // bucket offsets in entry order, 2 (nb + 1) per curve, so that
            // pass 1 finds a curve's from its place in the entry alone
            G.bkt = bkt_ints;
            bkt_ints += 2 * ((1 << G.p2) / G.c1 + 1);
            G.grid = (chunk > 0) ? acc : grid_total;
            acc += cells;
            grid_total += cells;
            grid_region = std::max(grid_region, acc);
            bcurves.push_back(ci);
            ++B.count;
        }
        P->buckets.push_back(B);
    }
    // grids larger than kStreamStoreBytes go out of pass 1 non-temporally:
    // they would not stay in the memory-side cache until pass 2 reads them,
    // and their write-back then costs pass 2 (measured: pass 2 of the survey
    // batch -4%); smaller ones stay cached for pass 2 (the long curve's 75 MB:
    // +2% when stored non-temporally)
    for (auto& B : P->buckets) {
        int64_t bytes = 0;
        for (int q = 0; q < B.count; ++q)
            bytes += (int64_t)ngrids * grid_cells(B.n1, B.p2) * (int64_t)sizeof(float2);
        B.nts = bytes > kStreamStoreBytes;
    }
    // pass 2 does not depend on the spreading mode: the float and the
    // fixed-point entries of one size bucket (adjacent in the curve list)
    // share a launch, unless grids are chunked (a chunk's pass 2 must run
    // before the next chunk's pass 1 reuses its region)
    for (size_t bi = 0; bi + 1 < P->buckets.size(); ++bi) {
        auto& B = P->buckets[bi];
        const auto& N = P->buckets[bi + 1];
        B.pass2_with_next = chunk <= 0 && N.n1 == B.n1 && N.p2 == B.p2 && N.cr == B.cr &&
                            N.first == B.first + B.count;
    }
    const int64_t grid_elems = (chunk > 0) ? grid_region : grid_total;
    P->off_nd0 = L.take(sizeof(int4) * total_pts);
    if (ngrids == 3) {
        P->off_nd1 = L.take(sizeof(int4) * total_pts);
        P->off_nd2 = L.take(sizeof(int4) * total_pts);
        // prepared points for the low-frequency direct sums
        P->off_w = L.take(sizeof(float) * total_pts);
        P->off_yw = L.take(sizeof(float) * total_pts);
        P->off_dt = L.take(sizeof(double) * total_pts);
    }
    P->off_bkt = L.take(sizeof(int32_t) * bkt_ints);
    P->nwide_aux = 0;
    P->wide_max_obs = 0;
    for (const CurveGeom& G : P->geom)
        if (G.n_obs > kWidePrep) {
            ++P->nwide_aux;
            P->wide_max_obs = std::max(P->wide_max_obs, (int)G.n_obs);
        }
    if (P->nwide_aux > 0) {
        P->off_wide = L.take(sizeof(WideAux) * (size_t)P->nwide_aux);
        const size_t nchunks = ((size_t)P->wide_max_obs + kWideChunk - 1) / kWideChunk;
        P->off_wpart = L.take(sizeof(WidePart) * nchunks * (size_t)P->nwide_aux);
    }
    P->off_grid = L.take(sizeof(float2) * grid_elems);
    P->bcurves = bcurves;
    // lanes: buckets own disjoint grid regions unless chunked, so they may run
    // on kAuxStreams + 1 streams at once; largest first onto the least loaded.
    // Entries that share a pass-2 launch form one group on one lane
    P->bucket_lane.assign(P->buckets.size(), 0);
    if (chunk <= 0 && P->buckets.size() > 1) {
        std::vector<std::pair<double, int>> order;   // (-cells, first entry)
        for (size_t i = 0; i < P->buckets.size(); ++i) {
            if (i > 0 && P->buckets[i - 1].pass2_with_next) continue;
            double cells = 0.0;
            for (size_t k = i; k < P->buckets.size(); ++k) {
                const auto& B = P->buckets[k];
                for (int q = 0; q < B.count; ++q)
                    cells += (double)P->geom[bcurves[B.first + q]].n;
                if (!B.pass2_with_next) break;
            }
            order.push_back({-cells, (int)i});
        }
        std::sort(order.begin(), order.end());
        double load[kAuxStreams + 1] = {};
        for (auto& o : order) {
            int best = 0;
            for (int j = 1; j <= kAuxStreams; ++j)
                if (load[j] < load[best]) best = j;
            load[best] += -o.first;
            for (size_t k = (size_t)o.second; k < P->buckets.size(); ++k) {
                P->bucket_lane[k] = best;
                if (!P->buckets[k].pass2_with_next) break;
            }
        }
    }
    return GLS_OK;
}

int plan_direct_tiles(gls_plan* P, std::vector<int32_t>& tiles) {
    int rf = (P->pp.method == GLS_METHOD_DIRECT_CUSTOM) ? 1 : 8;
    while (rf > 1) {
        int64_t nt = 0;
        for (auto& G : P->geom) nt += (G.nf + 256 * rf - 1) / (256 * rf);
        if (nt >= 512) break;
        rf >>= 1;
    }
    std::vector<int32_t> tc, tf;
    for (int i = 0; i < P->ncurves; ++i) {
        const int nf = P->geom[i].nf;
        for (int k = 0; k < nf; k += 256 * rf) {
            tc.push_back(i);
            tf.push_back(k);
        }
    }
    // one launch of 256-thread workgroups, one per tile
    if ((int64_t)tc.size() * 256 > kMaxLaunchItems)
        return fail(GLS_E_UNSUPPORTED, "too many direct-sum tiles for one launch: split the batch");
    P->rf = rf;
    P->ntiles = (int)tc.size();
    tiles = tc;
    tiles.insert(tiles.end(), tf.begin(), tf.end());
    return GLS_OK;
}

// Gauss-Legendre nodes and weights on [-1, 1] (Newton on P_n)
void gauss_legendre(int nq, std::vector<double>& x, std::vector<double>& w) {
    x.resize(nq);
    w.resize(nq);
    for (int i = 0; i < (nq + 1) / 2; ++i) {
        double z = std::cos(M_PI * (i + 0.75) / (nq + 0.5)), dp = 0.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = 0.0;
            for (int j = 1; j <= nq; ++j) {
                const double p2 = p1;
                p1 = p0;
                p0 = ((2.0 * j - 1.0) * z * p1 - (j - 1.0) * p2) / j;
            }
            dp = nq * (z * p0 - p1) / (z * z - 1.0);
            const double dz = p0 / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-15) break;
        }
        x[i] = -z;
        x[nq - 1 - i] = z;
        w[i] = w[nq - 1 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
    }
}

// 1/psi_hat(xi) of the ES window psi(x) = exp(beta (sqrt(1 - (x/a)^2) - 1)),
// a = m + 1/2 cells, at xi = i/(2 kEsTab), i = 0 .. kEsTab + 1:
// psi_hat(xi) = 2 int_0^a psi(x) cos(2 pi xi x) dx by Gauss-Legendre.
// Cached per (m, beta): plans are created per batch.
const std::vector<float>& es_deconv_table(int m, float beta) {
    static std::mutex mu;
    static std::map<std::pair<int, float>, std::vector<float>> cache;
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find({m, beta});
    if (it != cache.end()) return it->second;
    const int nq = 200;
    std::vector<double> gx, gw;
    gauss_legendre(nq, gx, gw);
    const double a = m + 0.5;
    std::vector<double> xs(nq), ws(nq);
    for (int q = 0; q < nq; ++q) {   // map [-1, 1] -> [0, a]
        const double x = 0.5 * a * (gx[q] + 1.0);
        const double z = x / a;
        xs[q] = x;
        ws[q] = 0.5 * a * gw[q] * std::exp((double)beta * (std::sqrt(std::max(0.0, 1.0 - z * z)) - 1.0));
    }
    std::vector<float> tab(kEsTab + 2);
    for (int i = 0; i <= kEsTab + 1; ++i) {
        const double xi = (double)i / (2.0 * kEsTab);
        double acc = 0.0;
        for (int q = 0; q < nq; ++q) acc += ws[q] * std::cos(2.0 * M_PI * xi * xs[q]);
        tab[i] = (float)(1.0 / (2.0 * acc));
    }
    return cache.emplace(std::make_pair(m, beta), std::move(tab)).first->second;
}

}  // namespace

int64_t max_launch_items() { return g_max_launch_items.load(std::memory_order_relaxed); }

int es_table_device(int m, float beta, const float** out) {
    static std::mutex mu;
    static std::map<std::pair<int, float>, float*> dev;
    const std::vector<float>& tab = es_deconv_table(m, beta);
    std::lock_guard<std::mutex> lk(mu);
    auto it = dev.find({m, beta});
    if (it == dev.end()) {
        float* p = nullptr;
        const size_t bytes = sizeof(float) * tab.size();
        GLS_HIP_TRY(hipMalloc((void**)&p, bytes));
        hipError_t e = hipMemcpy(p, tab.data(), bytes, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            (void)hipFree(p);
            return hip_fail(e, "hipMemcpy (ES table)");
        }
        it = dev.emplace(std::make_pair(m, beta), p).first;
    }
    *out = it->second;
    return GLS_OK;
}

namespace {

// Algorithmic bytes per stage: what each kernel must read and write at
// minimum for its role in the fused pipeline.
void build_stage_list(gls_plan* P) {
    P->stages.clear();
    const int ngr = (P->kind == 1) ? 1 : 3;
    double prep = 0.0;
    for (auto& G : P->geom) {
        const double in_bytes = (P->pp.input_kind == GLS_INPUT_RAW || P->kind == 1) ? 24.0 : 16.0;
        const double rec = (P->pp.method == GLS_METHOD_NFFT) ? 16.0 * ngr : 16.0;
        prep += (double)G.n_obs * (in_bytes + rec);
    }
    P->stages.push_back({0, 0, prep, -1});
    if (P->pp.method == GLS_METHOD_NFFT) {
        double b2 = 0.0;   // pass-2 bytes of the open launch
        int b2first = 0;   // its first entry
        for (size_t bi = 0; bi < P->buckets.size(); ++bi) {
            const auto& B = P->buckets[bi];
            if (bi == 0 || !P->buckets[bi - 1].pass2_with_next) {
                b2 = 0.0;
                b2first = (int)bi;
            }
            double b1 = 0.0;
            for (int q = 0; q < B.count; ++q) {
                const CurveGeom& G = P->geom[P->bcurves[B.first + q]];
                const double n = (double)G.n;
                b1 += ngr * (16.0 * G.n_obs + 8.0 * n);
                b2 += ngr * 8.0 * n + (P->kind == 1 ? 8.0 : 4.0) * G.nf;
            }
            P->stages.push_back({1, B.n1 << B.p2, b1, (int)bi});
            if (!B.pass2_with_next) P->stages.push_back({2, B.n1 << B.p2, b2, b2first});
        }
        if (P->kind == 0 && !P->pp.fp32_coords) {
            double bl = 0.0;
            for (auto& G : P->geom) bl += 16.0 * G.n_obs;
            P->stages.push_back({4, 0, bl, -1});
        }
    } else {
        double b = 0.0;
        for (auto& G : P->geom) b += 16.0 * G.n_obs + 4.0 * G.nf;
        P->stages.push_back({3, 0, b, -1});
    }
}

// lays out [geom | bucket curves | tiles] at the start of the workspace
void reserve_meta(gls_plan* P, Layout& L, size_t nbcurves, size_t ntiles_ints) {
    // prep order: curves up to kWidePrep points first, then the long ones
    P->prep_list.clear();
    for (int i = 0; i < P->ncurves; ++i)
        if (P->geom[i].n_obs <= kWidePrep) P->prep_list.push_back(i);
    P->nprep_small = (int)P->prep_list.size();
    for (int i = 0; i < P->ncurves; ++i)
        if (P->geom[i].n_obs > kWidePrep) P->prep_list.push_back(i);
    P->off_geom = L.take(sizeof(CurveGeom) * P->geom.size());
    P->off_prep = L.take(sizeof(int32_t) * P->prep_list.size());
    P->off_bgeom = L.take(sizeof(CurveGeom) * nbcurves);
    P->off_tiles = L.take(sizeof(int32_t) * ntiles_ints);
}

void fill_meta(gls_plan* P, const std::vector<int32_t>& bcurves,
               const std::vector<int32_t>& tiles) {
    P->meta.assign(P->off_tiles + sizeof(int32_t) * tiles.size(), 0);
    if (!P->prep_list.empty())
        std::memcpy(P->meta.data() + P->off_prep, P->prep_list.data(),
                    sizeof(int32_t) * P->prep_list.size());
    if (!P->geom.empty())
        std::memcpy(P->meta.data() + P->off_geom, P->geom.data(),
                    sizeof(CurveGeom) * P->geom.size());
    // the geometry again in bucket order, for the pass kernels
    for (size_t i = 0; i < bcurves.size(); ++i)
        std::memcpy(P->meta.data() + P->off_bgeom + i * sizeof(CurveGeom),
                    &P->geom[bcurves[i]], sizeof(CurveGeom));
    if (!tiles.empty())
        std::memcpy(P->meta.data() + P->off_tiles, tiles.data(), sizeof(int32_t) * tiles.size());
}

// rewrites the geometry regions of the host descriptors (curve order and
// bucket order) from P->geom; the prep list and tiles are unchanged
void refresh_meta_geometry(gls_plan* P) {
    if (!P->geom.empty())
        std::memcpy(P->meta.data() + P->off_geom, P->geom.data(),
                    sizeof(CurveGeom) * P->geom.size());
    for (size_t i = 0; i < P->bcurves.size(); ++i)
        std::memcpy(P->meta.data() + P->off_bgeom + i * sizeof(CurveGeom),
                    &P->geom[P->bcurves[i]], sizeof(CurveGeom));
}

void destroy_plan(gls_plan* P) {
    if (!P) return;
    for (auto e : P->events) (void)hipEventDestroy(e);
    for (auto e : P->fork_join) (void)hipEventDestroy(e);
    delete P;
}

PointArrays point_arrays(const gls_plan* P, void* ws) {
    char* base = (char*)ws;
    PointArrays pa{};
    pa.nd0 = (int4*)(base + P->off_nd0);
    pa.nd1 = (int4*)(base + P->off_nd1);
    pa.nd2 = (int4*)(base + P->off_nd2);
    // offset 0 is the descriptor region: it marks an array the plan lacks
    pa.w = P->off_w ? (float*)(base + P->off_w) : nullptr;
    pa.yw = P->off_yw ? (float*)(base + P->off_yw) : nullptr;
    pa.dt = P->off_dt ? (double*)(base + P->off_dt) : nullptr;
    return pa;
}

// The whole launch sequence of a plan on stream s.
int run_plan(gls_plan* P, const double* t, const void* a, const void* b, void* out,
             int out_f64, int y_is_f32, void* ws, hipStream_t s, const double* freqs = nullptr,
             bool upload_meta = true, bool allow_fork = true) {
    char* base = (char*)ws;
    if (upload_meta && !P->meta.empty())
        GLS_HIP_TRY(upload_async(base, P->meta.data(), P->meta.size(), s));
    const CurveGeom* d_geom = (const CurveGeom*)(base + P->off_geom);
    const CurveGeom* d_bgeom = (const CurveGeom*)(base + P->off_bgeom);
    const int32_t* d_tiles = (const int32_t*)(base + P->off_tiles);
    CurveState* state = (CurveState*)(base + P->off_state);
    PointArrays pa = point_arrays(P, ws);
    int32_t* bkt = (int32_t*)(base + P->off_bkt);
    float2* grids = (float2*)(base + P->off_grid);
    int ev = 0;
    auto mark = [&]() -> hipError_t {
        if (!P->timing) return hipSuccess;
        return hipEventRecord(P->events[ev++], s);
    };
    const int ngrids = (P->kind == 1) ? 1 : 3;
    const PlanParams& pp = P->pp;
    GLS_HIP_TRY(mark());
    const int32_t* d_prep = (const int32_t*)(base + P->off_prep);
    WideAux* aux = P->nwide_aux > 0 ? (WideAux*)(base + P->off_wide) : nullptr;
    WidePart* wpart = P->nwide_aux > 0 ? (WidePart*)(base + P->off_wpart) : nullptr;
    GLS_HIP_TRY(launch_prep(d_geom, d_prep, P->nprep_small, d_prep + P->nprep_small,
                            P->ncurves - P->nprep_small, state, t, a, b, pa, bkt, pp, y_is_f32,
                            aux, wpart, P->wide_max_obs, s));
    GLS_HIP_TRY(mark());
    if (P->pp.method == GLS_METHOD_NFFT) {
        // fork the buckets over the caller's and the auxiliary streams (not
        // when timing, whose events need one serial stream), join at the end
        bool fork = false;
        for (int lane : P->bucket_lane) fork |= lane > 0;
        fork &= P->fork_buckets && !P->timing && allow_fork;
        hipStream_t lanes[kAuxStreams + 1] = {s};
        if (fork) {
            GLS_HIP_TRY(aux_streams(lanes + 1));
            if (P->fork_join.empty()) {
                P->fork_join.resize(kAuxStreams + 1);
                for (auto& e : P->fork_join)
                    GLS_HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            }
            GLS_HIP_TRY(hipEventRecord(P->fork_join[0], s));
            for (int j = 1; j <= kAuxStreams; ++j)
                GLS_HIP_TRY(hipStreamWaitEvent(lanes[j], P->fork_join[0], 0));
        }
        size_t gfirst = 0;   // first entry of the open pass-2 launch
        for (size_t bi = 0; bi < P->buckets.size(); ++bi) {
            const auto& B = P->buckets[bi];
            if (bi == 0 || !P->buckets[bi - 1].pass2_with_next) gfirst = bi;
            const hipStream_t ls = fork ? lanes[P->bucket_lane[bi]] : s;
            GLS_HIP_TRY(launch_pass1(B.n1, B.c1, B.fix, B.nts, d_bgeom + B.first, B.count, ngrids,
                                     B.p2, pa, bkt + B.bkt0, grids, pp, ls));
            GLS_HIP_TRY(mark());
            if (B.pass2_with_next) continue;
            // pass 2 over this entry and the ones whose pass 2 it takes
            const auto& F = P->buckets[gfirst];
            const CurveGeom* g2 = d_bgeom + F.first;
            const int n2curves = (int)(B.first + B.count - F.first);
            if (P->kind == 1)
                GLS_HIP_TRY(launch_pass2_nfft(F.p2, F.cr, F.n1, g2, state, n2curves, grids, pp,
                                              (float2*)out, ls));
            else
                GLS_HIP_TRY(launch_pass2_lomb(F.p2, F.cr, F.n1, g2, state, n2curves, grids, pp,
                                              out, out_f64, ls));
            GLS_HIP_TRY(mark());
        }
        if (fork) {
            for (int j = 1; j <= kAuxStreams; ++j) {
                GLS_HIP_TRY(hipEventRecord(P->fork_join[j], lanes[j]));
                GLS_HIP_TRY(hipStreamWaitEvent(s, P->fork_join[j], 0));
            }
        }
        if (P->kind == 0 && !P->pp.fp32_coords) {
            GLS_HIP_TRY(launch_lowfreq(d_geom, d_prep, P->nprep_small, d_prep + P->nprep_small,
                                       P->ncurves - P->nprep_small, state, pa, pp, out,
                                       out_f64, aux, s));
            GLS_HIP_TRY(mark());
        }
    } else if (P->pp.method == GLS_METHOD_DIRECT_CUSTOM) {
        GLS_HIP_TRY(launch_dirsum_custom(d_geom, state, d_tiles, d_tiles + P->ntiles, P->ntiles,
                                         pa, P->pp, freqs, out, out_f64, s));
        GLS_HIP_TRY(mark());
    } else {
        GLS_HIP_TRY(launch_dirsum(d_geom, state, d_tiles, d_tiles + P->ntiles, P->ntiles, P->rf,
                                  pa, P->pp, out, out_f64, s));
        GLS_HIP_TRY(mark());
    }
    return GLS_OK;
}

int execute_lomb(gls_plan* P, const double* t, const void* a, const void* b, void* power,
                 int out_f64, void* ws, void* stream, bool allow_fork = true) {
    int rc = require_init();
    if (rc) return rc;
    if (!P) return fail(GLS_E_INVALID, "plan is NULL");
    if (P->kind != 0) return fail(GLS_E_INVALID, "plan is not a Lomb-Scargle plan");
    if (P->ncurves == 0) return GLS_OK;
    if (!t || !a || !b || !power || !ws) return fail(GLS_E_INVALID, "NULL device pointer");
    if (P->pp.method == GLS_METHOD_DIRECT_CUSTOM)
        return fail(GLS_E_INVALID, "custom-frequency plans run with gls_lomb_execute_custom");
    return run_plan(P, t, a, b, power, out_f64, 0, ws, (hipStream_t)stream, nullptr, true,
                    allow_fork);
}

int finish_plan(gls_plan* P, Layout& L) {
    build_stage_list(P);
    P->ws_bytes = std::max<size_t>(L.cur, 256);
    return GLS_OK;
}

int check_curve(int64_t obs, int32_t n_obs, int32_t nf, double f0, double df, int64_t out) {
    return (n_obs < 0 || nf < 1 || obs < 0 || out < 0 || !std::isfinite(f0) ||
            !std::isfinite(df) || !(df > 0.0))
               ? 1
               : 0;
}

}  // namespace
}  // namespace gls

using namespace gls;

extern "C" {

int gls_lomb_plan_create(gls_plan** plan, const gls_curve* curves, int32_t ncurves,
                         const gls_params* params) {
    int rc = require_init();
    if (rc) return rc;
    if (!plan || !params) return fail(GLS_E_INVALID, "plan or params is NULL");
    *plan = nullptr;
    if (ncurves < 0 || (ncurves > 0 && !curves)) return fail(GLS_E_INVALID, "bad curve list");
    if (ncurves > kMaxPlanCurves)
        return fail(GLS_E_UNSUPPORTED, "more than 2^22 curves in one plan: split the batch");
    const gls_params& prm = *params;
    if (prm.mode < 0 || prm.mode > 2) return fail(GLS_E_INVALID, "mode must be 0, 1 or 2");
    if (prm.input_kind != GLS_INPUT_RAW && prm.input_kind != GLS_INPUT_WEIGHTED)
        return fail(GLS_E_INVALID, "unknown input_kind");
    if (prm.method != GLS_METHOD_NFFT && prm.method != GLS_METHOD_DIRECT &&
        prm.method != GLS_METHOD_DIRECT_CUSTOM)
        return fail(GLS_E_INVALID, "unknown method");
    if (prm.method == GLS_METHOD_NFFT) {
        if (!(prm.sigma >= 1.0 && prm.sigma <= 64.0))
            return fail(GLS_E_INVALID, "sigma must be in [1, 64]");
        if (prm.m < 1 || prm.m > 24) return fail(GLS_E_INVALID, "m must be in [1, 24]");
    }
    if (!std::isfinite(prm.reg_c) || !std::isfinite(prm.reg_s) || prm.reg_c < 0 || prm.reg_s < 0)
        return fail(GLS_E_INVALID, "regularisation must be finite and >= 0");
    if (prm.window != GLS_WINDOW_GAUSSIAN && prm.window != GLS_WINDOW_ES)
        return fail(GLS_E_INVALID, "unknown window");
    if (!std::isfinite(prm.window_beta) || prm.window_beta > 200.f)
        return fail(GLS_E_INVALID, "window_beta must be finite and <= 200");
    if (prm.point_capacity > ((int64_t)1 << 34))
        return fail(GLS_E_INVALID, "point_capacity above 2^34");
    if (prm.precision != GLS_PRECISION_FP64_COORDS && prm.precision != GLS_PRECISION_FP32)
        return fail(GLS_E_INVALID, "unknown precision");
    if (prm.precision == GLS_PRECISION_FP32 && prm.method != GLS_METHOD_NFFT)
        return fail(GLS_E_INVALID, "GLS_PRECISION_FP32 is an NFFT-method option");
    if (prm.spread != GLS_SPREAD_FLOAT && prm.spread != GLS_SPREAD_FIXED &&
        prm.spread != GLS_SPREAD_AUTO)
        return fail(GLS_E_INVALID, "unknown spread mode");
    gls_plan* P = new gls_plan();
    P->id = g_plan_ids.fetch_add(1);
    P->kind = 0;
    P->pp.m = prm.m;
    P->pp.mode = prm.mode;
    P->pp.reg_c = prm.reg_c;
    P->pp.reg_s = prm.reg_s;
    P->pp.input_kind = prm.input_kind;
    P->pp.method = prm.method;
    P->pp.kind = 0;
    P->pp.window = (prm.method == GLS_METHOD_NFFT) ? prm.window : GLS_WINDOW_GAUSSIAN;
    P->pp.fp32_coords = prm.precision == GLS_PRECISION_FP32 ? 1 : 0;
    P->spread_mode = prm.spread;
    P->pp.es_beta = prm.window_beta > 0.f
                        ? prm.window_beta
                        : (float)(M_PI * (2 * prm.m + 1) * (1.0 - 0.5 / prm.sigma));
    P->pp.es_deconv = nullptr;
    if (P->pp.window == GLS_WINDOW_ES) {
        rc = es_table_device(prm.m, P->pp.es_beta, &P->pp.es_deconv);
        if (rc != GLS_OK) {
            destroy_plan(P);
            return rc;
        }
    }
    P->sigma = prm.sigma;
    P->chunk_bytes = std::max<int64_t>(prm.grid_chunk_bytes, 0);
    P->ncurves = ncurves;
    P->geom.resize(ncurves);
    int64_t pts = 0;
    for (int i = 0; i < ncurves; ++i) {
        const gls_curve& c = curves[i];
        const bool custom = prm.method == GLS_METHOD_DIRECT_CUSTOM;
        if (check_curve(c.obs_offset, c.n_obs, c.nf, custom ? 0.0 : c.f0, custom ? 1.0 : c.df,
                        c.out_offset)) {
            destroy_plan(P);
            return fail(GLS_E_INVALID, "invalid curve descriptor " + std::to_string(i));
        }
        CurveGeom& G = P->geom[i];
        G = CurveGeom{};
        G.ci = i;
        G.obs = c.obs_offset;
        G.out = c.out_offset;
        G.pts = pts;
        G.f0 = c.f0;
        G.df = c.df;
        G.ybar_in = c.ybar;
        G.yy_in = c.yy;
        G.n_obs = c.n_obs;
        G.nf = c.nf;
        pts += c.n_obs;
        P->total_out = std::max<int64_t>(P->total_out, c.out_offset + c.nf);
    }
    P->pts_cap = std::max<int64_t>(pts, prm.point_capacity);
    std::vector<int32_t> bcurves, tiles;
    if (prm.method != GLS_METHOD_NFFT) rc = plan_direct_tiles(P, tiles);
    if (rc != GLS_OK) {
        destroy_plan(P);
        return rc;
    }
    // bucket curve lists need the geometry first: plan into a scratch layout,
    // then lay out [meta | state | data] for real
    Layout L;
    if (prm.method == GLS_METHOD_NFFT) {
        Layout probe;
        rc = plan_nfft_geometry(P, 3, probe, P->pts_cap, bcurves);
        if (rc != GLS_OK) {
            destroy_plan(P);
            return rc;
        }
        reserve_meta(P, L, bcurves.size(), 0);
        P->off_state = L.take(sizeof(CurveState) * (size_t)ncurves);
        // re-take the data regions after the meta and state regions
        const size_t shift = L.cur;
        for (size_t* o : {&P->off_nd0, &P->off_nd1, &P->off_nd2, &P->off_w, &P->off_yw,
                          &P->off_dt,
                          &P->off_bkt, &P->off_wide, &P->off_wpart, &P->off_grid})
            *o += shift;
        L.cur += probe.cur;
    } else {
        reserve_meta(P, L, 0, tiles.size());
        P->off_state = L.take(sizeof(CurveState) * (size_t)ncurves);
        P->off_w = L.take(sizeof(float) * P->pts_cap);
        P->off_yw = L.take(sizeof(float) * P->pts_cap);
        P->off_dt = L.take(sizeof(double) * P->pts_cap);
    }
    fill_meta(P, bcurves, tiles);
    rc = finish_plan(P, L);
    if (rc != GLS_OK) {
        destroy_plan(P);
        return rc;
    }
    *plan = P;
    return GLS_OK;
}

int gls_nfft_plan_create(gls_plan** plan, const gls_nfft_item* items, int32_t nitems,
                         double sigma, int32_t m) {
    int rc = require_init();
    if (rc) return rc;
    if (!plan) return fail(GLS_E_INVALID, "plan is NULL");
    *plan = nullptr;
    if (nitems < 0 || (nitems > 0 && !items)) return fail(GLS_E_INVALID, "bad item list");
    if (nitems > kMaxPlanCurves)
        return fail(GLS_E_UNSUPPORTED, "more than 2^22 items in one plan: split the batch");
    if (!(sigma >= 1.0 && sigma <= 64.0)) return fail(GLS_E_INVALID, "sigma must be in [1, 64]");
    if (m < 1 || m > 24) return fail(GLS_E_INVALID, "m must be in [1, 24]");
    gls_plan* P = new gls_plan();
    P->id = g_plan_ids.fetch_add(1);
    P->kind = 1;
    P->pp.m = m;
    P->pp.mode = GLS_MODE_FLOATING_MEAN;
    P->pp.method = GLS_METHOD_NFFT;
    P->pp.kind = 1;
    P->sigma = sigma;
    P->chunk_bytes = 0;
    P->ncurves = nitems;
    P->geom.resize(nitems);
    int64_t pts = 0;
    for (int i = 0; i < nitems; ++i) {
        const gls_nfft_item& c = items[i];
        if (check_curve(c.obs_offset, c.n_obs, c.nf, c.f0, c.df, c.out_offset)) {
            destroy_plan(P);
            return fail(GLS_E_INVALID, "invalid NFFT item " + std::to_string(i));
        }
        CurveGeom& G = P->geom[i];
        G = CurveGeom{};
        G.ci = i;
        G.obs = c.obs_offset;
        G.out = c.out_offset;
        G.pts = pts;
        G.f0 = c.f0;
        G.df = c.df;
        G.n_obs = c.n_obs;
        G.nf = c.nf;
        pts += c.n_obs;
        P->total_out = std::max<int64_t>(P->total_out, c.out_offset + c.nf);
    }
    P->pts_cap = pts;
    std::vector<int32_t> bcurves, tiles;
    Layout probe, L;
    rc = plan_nfft_geometry(P, 1, probe, P->pts_cap, bcurves);
    if (rc != GLS_OK) {
        destroy_plan(P);
        return rc;
    }
    reserve_meta(P, L, bcurves.size(), 0);
    P->off_state = L.take(sizeof(CurveState) * (size_t)nitems);
    const size_t shift = L.cur;
    for (size_t* o : {&P->off_nd0, &P->off_bkt, &P->off_wide,
                      &P->off_wpart, &P->off_grid})
        *o += shift;
    L.cur += probe.cur;
    fill_meta(P, bcurves, tiles);
    rc = finish_plan(P, L);
    if (rc != GLS_OK) {
        destroy_plan(P);
        return rc;
    }
    *plan = P;
    return GLS_OK;
}

int gls_plan_destroy(gls_plan* plan) {
    destroy_plan(plan);
    return GLS_OK;
}

int gls_plan_workspace_bytes(const gls_plan* plan, size_t* bytes) {
    if (!plan || !bytes) return fail(GLS_E_INVALID, "NULL argument");
    *bytes = plan->ws_bytes;
    return GLS_OK;
}

int gls_plan_output_count(const gls_plan* plan, int64_t* count) {
    if (!plan || !count) return fail(GLS_E_INVALID, "NULL argument");
    *count = plan->total_out;
    return GLS_OK;
}

int gls_plan_grid_size(const gls_plan* plan, int32_t i, int64_t* n) {
    if (!plan || !n) return fail(GLS_E_INVALID, "NULL argument");
    if (i < 0 || i >= plan->ncurves) return fail(GLS_E_INVALID, "curve index out of range");
    *n = (plan->pp.method == GLS_METHOD_NFFT) ? (int64_t)plan->geom[i].n : 0;
    return GLS_OK;
}

int gls_plan_curve_geometry(const gls_plan* plan, int32_t i, int64_t* n, int32_t* n1,
                            int32_t* n2, double* b) {
    if (!plan) return fail(GLS_E_INVALID, "plan is NULL");
    if (i < 0 || i >= plan->ncurves) return fail(GLS_E_INVALID, "curve index out of range");
    const CurveGeom& G = plan->geom[i];
    const bool nfft = plan->pp.method == GLS_METHOD_NFFT;
    if (n) *n = nfft ? (int64_t)G.n : 0;
    if (n1) *n1 = nfft ? G.n1 : 0;
    if (n2) *n2 = nfft ? (1 << G.p2) : 0;
    if (b) *b = nfft ? G.b : 0.0;
    return GLS_OK;
}

int gls_plan_curve_tiles(const gls_plan* plan, int32_t i, int32_t* c1, int32_t* cr) {
    if (!plan) return fail(GLS_E_INVALID, "plan is NULL");
    if (i < 0 || i >= plan->ncurves) return fail(GLS_E_INVALID, "curve index out of range");
    const CurveGeom& G = plan->geom[i];
    int tc = 0, tr = 0;
    if (plan->pp.method == GLS_METHOD_NFFT)
        for (const auto& B : plan->buckets)
            if (B.n1 == G.n1 && B.p2 == G.p2 && B.fix == (plan->curve_fixed[i] != 0)) {
                tc = B.c1;
                tr = B.cr;
                break;
            }
    if (c1) *c1 = tc;
    if (cr) *cr = tr;
    return GLS_OK;
}

int gls_plan_curve_spread(const gls_plan* plan, int32_t i, int32_t* spread) {
    if (!plan) return fail(GLS_E_INVALID, "plan is NULL");
    if (i < 0 || i >= plan->ncurves) return fail(GLS_E_INVALID, "curve index out of range");
    if (!spread) return fail(GLS_E_INVALID, "NULL output");
    const bool fix = plan->pp.method == GLS_METHOD_NFFT && i < (int)plan->curve_fixed.size() &&
                     plan->curve_fixed[i] != 0;
    *spread = fix ? GLS_SPREAD_FIXED : GLS_SPREAD_FLOAT;
    return GLS_OK;
}

int gls_tile_set(int32_t n1, int32_t p2, int32_t c1, int32_t cr) {
    bool in_catalogue = false;
    for (int v : kPass1Sizes) in_catalogue |= v == n1;
    if (!in_catalogue || p2 < kMinLog2Row || p2 > kMaxLog2Row)
        return fail(GLS_E_INVALID, "(n1, p2) is not a grid size of the catalogue");
    if (c1 < 0 || cr < 0) return fail(GLS_E_INVALID, "tiles must be >= 0");
    if (c1 && !(pass1_tile_ok(n1, c1) && pass1_tile_built(n1, c1)))
        return fail(GLS_E_UNSUPPORTED, "no pass-1 kernel with " + std::to_string(c1) +
                                           " columns of " + std::to_string(n1) +
                                           " cells in this library");
    if (cr && !(pass2_tile_ok(p2, cr) && pass2_tile_built(p2, cr)))
        return fail(GLS_E_UNSUPPORTED, "no pass-2 kernel with " + std::to_string(cr) +
                                           " rows of 2^" + std::to_string(p2) +
                                           " cells in this library");
    std::lock_guard<std::mutex> lk(g_tile_mu);
    if (!c1 && !cr) g_tile_override.erase({n1, p2});
    else g_tile_override[{n1, p2}] = BucketTiles{c1, cr};
    return GLS_OK;
}

int gls_launch_items_set(int64_t items) {
    if (items > kMaxLaunchItems)
        return fail(GLS_E_INVALID, "at most 2^31 work-items per launch");
    g_max_launch_items.store(items > 0 ? items : kMaxLaunchItems);
    return GLS_OK;
}

int gls_tile_get(int32_t n1, int32_t p2, int32_t* c1, int32_t* cr) {
    const BucketTiles t = current_tiles(n1, p2);
    if (c1) *c1 = t.c1;
    if (cr) *cr = t.cr;
    return GLS_OK;
}

int gls_grid_cost_set(int32_t n1, int32_t p2, double ps_per_cell) {
    bool in_catalogue = false;
    for (int v : kPass1Sizes) in_catalogue |= v == n1;
    if (!in_catalogue || p2 < kMinLog2Row || p2 > kMaxLog2Row)
        return fail(GLS_E_INVALID, "(n1, p2) is not a grid size of the catalogue");
    if (!std::isfinite(ps_per_cell)) return fail(GLS_E_INVALID, "cost must be finite");
    std::lock_guard<std::mutex> lk(g_tile_mu);
    if (ps_per_cell <= 0) g_cost_override.erase({n1, p2});
    else g_cost_override[{n1, p2}] = ps_per_cell;
    return GLS_OK;
}

int gls_grid_cost_get(int32_t n1, int32_t p2, double* ps_per_cell, int32_t* source) {
    const int i = size_index(n1);
    if (i < 0 || p2 < kMinLog2Row || p2 > kMaxLog2Row)
        return fail(GLS_E_INVALID, "(n1, p2) is not a grid size of the catalogue");
    double c = table_index().ps[p2 - kMinLog2Row][i];
    int src = table_index().measured[p2 - kMinLog2Row][i] ? 1 : 0;
    {
        std::lock_guard<std::mutex> lk(g_tile_mu);
        auto it = g_cost_override.find({n1, p2});
        if (it != g_cost_override.end()) {
            c = it->second;
            src = 2;
        }
    }
    if (ps_per_cell) *ps_per_cell = c;
    if (source) *source = src;
    return GLS_OK;
}

int gls_lomb_execute(gls_plan* plan, const double* t, const void* a, const void* b, float* power,
                     void* workspace, void* stream) {
    return execute_lomb(plan, t, a, b, power, 0, workspace, stream);
}

int gls_lomb_execute_static(gls_plan* plan, const double* t, const void* a, const void* b,
                            float* power, void* stream) {
    int rc = require_init();
    if (rc) return rc;
    if (!plan) return fail(GLS_E_INVALID, "plan is NULL");
    if (plan->ws_bytes > kScratchBytes)
        return fail(GLS_E_UNSUPPORTED,
                    "plan needs " + std::to_string(plan->ws_bytes) +
                        " bytes of workspace; the static scratch of gls_lomb_execute_static "
                        "holds " + std::to_string(kScratchBytes) +
                        " (run this plan with gls_lomb_execute and a workspace)");
    void* ws = nullptr;
    rc = static_scratch(&ws);
    if (rc) return rc;
    // buckets in series on the caller's stream (forking would create the
    // auxiliary streams and events on first use)
    return execute_lomb(plan, t, a, b, power, 0, ws, stream, false);
}

int gls_lomb_execute_custom(gls_plan* plan, const double* t, const void* a, const void* b,
                            const double* freqs, void* power, int32_t power_is_f64,
                            void* workspace, void* stream) {
    int rc = require_init();
    if (rc) return rc;
    if (!plan) return fail(GLS_E_INVALID, "plan is NULL");
    if (plan->kind != 0 || plan->pp.method != GLS_METHOD_DIRECT_CUSTOM)
        return fail(GLS_E_INVALID, "plan was not created with GLS_METHOD_DIRECT_CUSTOM");
    if (plan->ncurves == 0) return GLS_OK;
    if (!t || !a || !b || !freqs || !power || !workspace)
        return fail(GLS_E_INVALID, "NULL device pointer");
    return run_plan(plan, t, a, b, power, power_is_f64 ? 1 : 0, 0, workspace,
                    (hipStream_t)stream, freqs);
}

int gls_lomb_execute_f64(gls_plan* plan, const double* t, const void* a, const void* b,
                         double* power, void* workspace, void* stream) {
    return execute_lomb(plan, t, a, b, power, 1, workspace, stream);
}

int gls_nfft_execute(gls_plan* plan, const double* t, const void* y, int32_t y_is_f32,
                     float* out, void* workspace, void* stream) {
    int rc = require_init();
    if (rc) return rc;
    if (!plan) return fail(GLS_E_INVALID, "plan is NULL");
    if (plan->kind != 1) return fail(GLS_E_INVALID, "plan is not an NFFT plan");
    if (plan->ncurves == 0) return GLS_OK;
    if (!t || !y || !out || !workspace) return fail(GLS_E_INVALID, "NULL device pointer");
    return run_plan(plan, t, y, y, out, 0, y_is_f32 ? 1 : 0, workspace, (hipStream_t)stream);
}

int gls_gaussian_grid(const double* t, const double* y, int32_t n_obs, int32_t ng, int32_t m,
                      double b, double tmin, double tmax, double spp, float* grid,
                      void* stream) {
    int rc = require_init();
    if (rc) return rc;
    if (ng < 1 || n_obs < 0 || m < 0 || !(b > 0) || !grid || (n_obs > 0 && (!t || !y)))
        return fail(GLS_E_INVALID, "invalid gridding arguments");
    if (!(tmax > tmin) || !(spp > 0)) return fail(GLS_E_INVALID, "need tmax > tmin and spp > 0");
    GLS_HIP_TRY(launch_gaussian_grid(t, y, n_obs, ng, m, b, tmin, tmax, spp, grid,
                                     (hipStream_t)stream));
    return GLS_OK;
}

int gls_plan_set_fork(gls_plan* plan, int32_t enable) {
    if (!plan) return fail(GLS_E_INVALID, "plan is NULL");
    plan->fork_buckets = enable != 0;
    return GLS_OK;
}

int gls_plan_enable_timing(gls_plan* plan, int32_t enable) {
    if (!plan) return fail(GLS_E_INVALID, "plan is NULL");
    if (enable && plan->events.empty()) {
        plan->events.resize(plan->stages.size() + 1);
        for (auto& e : plan->events) GLS_HIP_TRY(hipEventCreate(&e));
    }
    plan->timing = enable != 0;
    return GLS_OK;
}

int gls_plan_stage_count(const gls_plan* plan, int32_t* count) {
    if (!plan || !count) return fail(GLS_E_INVALID, "NULL argument");
    *count = (int32_t)plan->stages.size();
    return GLS_OK;
}

int gls_plan_stage_bucket(const gls_plan* plan, int32_t stage, int32_t* n1, int32_t* p2,
                          int32_t* c1, int32_t* cr, int32_t* spread, int32_t* streamed) {
    if (!plan) return fail(GLS_E_INVALID, "plan is NULL");
    if (stage < 0 || stage >= (int)plan->stages.size())
        return fail(GLS_E_INVALID, "stage index out of range");
    const int b = plan->stages[stage].bucket;
    const gls_plan::Bucket* B = b >= 0 ? &plan->buckets[b] : nullptr;
    if (n1) *n1 = B ? B->n1 : 0;
    if (p2) *p2 = B ? B->p2 : 0;
    if (c1) *c1 = B ? B->c1 : 0;
    if (cr) *cr = B ? B->cr : 0;
    if (spread) *spread = (B && B->fix) ? GLS_SPREAD_FIXED : GLS_SPREAD_FLOAT;
    if (streamed) *streamed = (B && B->nts) ? 1 : 0;
    return GLS_OK;
}

int gls_plan_stage_times(const gls_plan* plan, int32_t* kind, int32_t* grid, float* ms,
                         double* bytes, int32_t max) {
    if (!plan) return fail(GLS_E_INVALID, "plan is NULL");
    const int ns = (int)plan->stages.size();
    for (int i = 0; i < ns && i < max; ++i) {
        if (kind) kind[i] = plan->stages[i].kind;
        if (grid) grid[i] = plan->stages[i].grid;
        if (bytes) bytes[i] = plan->stages[i].bytes;
        if (ms) {
            ms[i] = 0.f;
            if (plan->timing && (size_t)i + 1 < plan->events.size())
                GLS_HIP_TRY(hipEventElapsedTime(&ms[i], plan->events[i], plan->events[i + 1]));
        }
    }
    return GLS_OK;
}

int gls_topk(const gls_plan* plan, const void* workspace, const float* power,
             const uint8_t* mask, int32_t k, float* top_power, int32_t* top_index,
             void* stream) {
    int rc = require_init();
    if (rc) return rc;
    if (!plan || !workspace || !power || !top_power || !top_index)
        return fail(GLS_E_INVALID, "NULL argument");
    if (k < 1 || k > 64) return fail(GLS_E_INVALID, "k must be in [1, 64]");
    const CurveGeom* d_geom = (const CurveGeom*)((const char*)workspace + plan->off_geom);
    GLS_HIP_TRY(launch_topk(d_geom, plan->ncurves, power, mask, k, top_power, top_index,
                            (hipStream_t)stream));
    return GLS_OK;
}

// ---- captured executes ------------------------------------------------------
struct gls_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t plan_id = 0;     // the plan whose descriptors the graph's kernels read
    void* workspace = nullptr;
    size_t meta_bytes = 0;
};

int gls_plan_graph_create(gls_plan* plan, const double* t, const void* a, const void* b,
                          void* power, int32_t power_is_f64, void* workspace, void* stream,
                          gls_graph** out) {
    int rc = require_init();
    if (rc) return rc;
    if (!plan || !out) return fail(GLS_E_INVALID, "plan or out is NULL");
    *out = nullptr;
    if (plan->kind != 0 || plan->pp.method == GLS_METHOD_DIRECT_CUSTOM)
        return fail(GLS_E_UNSUPPORTED, "graphs are made of Lomb-Scargle NFFT or direct plans");
    if (plan->ncurves == 0) return fail(GLS_E_INVALID, "empty plan");
    if (!t || !a || !b || !power || !workspace) return fail(GLS_E_INVALID, "NULL device pointer");
    const hipStream_t s = (hipStream_t)stream;
    if (!s) return fail(GLS_E_INVALID, "capture needs a created stream (not the null stream)");
    // descriptors go into the workspace once, outside the graph
    if (!plan->meta.empty()) GLS_HIP_TRY(upload_async(workspace, plan->meta.data(),
                                                      plan->meta.size(), s));
    GLS_HIP_TRY(hipStreamSynchronize(s));
    const bool timing = plan->timing, fork = plan->fork_buckets;
    plan->timing = false;
    plan->fork_buckets = false;
    gls_graph* g = new gls_graph();
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        rc = run_plan(plan, t, a, b, power, power_is_f64 ? 1 : 0, 0, workspace, s, nullptr,
                      false);
        hipError_t e2 = hipStreamEndCapture(s, &g->graph);
        if (rc == GLS_OK && e2 != hipSuccess) rc = hip_fail(e2, "hipStreamEndCapture");
    } else {
        rc = hip_fail(e, "hipStreamBeginCapture");
    }
    plan->timing = timing;
    plan->fork_buckets = fork;
    if (rc == GLS_OK) {
        e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) rc = hip_fail(e, "hipGraphInstantiate");
    }
    if (rc != GLS_OK) {
        if (g->exec) (void)hipGraphExecDestroy(g->exec);
        if (g->graph) (void)hipGraphDestroy(g->graph);
        delete g;
        return rc;
    }
    g->plan_id = plan->id;
    g->workspace = workspace;
    g->meta_bytes = plan->meta.size();
    *out = g;
    return GLS_OK;
}

int gls_plan_rebind(gls_plan* plan, const gls_curve* curves, int32_t ncurves) {
    if (!plan) return fail(GLS_E_INVALID, "plan is NULL");
    if (plan->kind != 0) return fail(GLS_E_INVALID, "plan is not a Lomb-Scargle plan");
    if (ncurves != plan->ncurves)
        return fail(GLS_E_INVALID, "rebind needs as many curves as the plan (" +
                                       std::to_string(plan->ncurves) + ")");
    if (ncurves > 0 && !curves) return fail(GLS_E_INVALID, "curves is NULL");
    const bool custom = plan->pp.method == GLS_METHOD_DIRECT_CUSTOM;
    int64_t pts = 0;
    int wide_max = 0;
    for (int i = 0; i < ncurves; ++i) {
        const gls_curve& c = curves[i];
        const CurveGeom& G = plan->geom[i];
        if (check_curve(c.obs_offset, c.n_obs, c.nf, custom ? 0.0 : c.f0, custom ? 1.0 : c.df,
                        c.out_offset))
            return fail(GLS_E_INVALID, "invalid curve descriptor " + std::to_string(i));
        if (c.nf != G.nf || (!custom && (c.f0 != G.f0 || c.df != G.df)))
            return fail(GLS_E_INVALID, "curve " + std::to_string(i) +
                                           ": a rebind keeps every curve's frequency grid");
        if ((c.n_obs > kWidePrep) != (G.n_obs > kWidePrep))
            return fail(GLS_E_INVALID, "curve " + std::to_string(i) +
                                           " crosses the 16,384-point preparation limit");
        if (c.n_obs > kWidePrep) wide_max = std::max(wide_max, (int)c.n_obs);
        pts += c.n_obs;
    }
    if (pts > plan->pts_cap)
        return fail(GLS_E_INVALID, "rebind of " + std::to_string(pts) +
                                       " points exceeds the plan's point capacity " +
                                       std::to_string(plan->pts_cap));
    if (wide_max > plan->wide_max_obs)
        return fail(GLS_E_INVALID, "a curve above 16,384 points is longer than the plan's");
    int64_t off = 0, total_out = 0;
    for (int i = 0; i < ncurves; ++i) {
        const gls_curve& c = curves[i];
        CurveGeom& G = plan->geom[i];
        G.obs = c.obs_offset;
        G.n_obs = c.n_obs;
        G.pts = off;
        G.out = c.out_offset;
        G.ybar_in = c.ybar;
        G.yy_in = c.yy;
        off += c.n_obs;
        total_out = std::max<int64_t>(total_out, c.out_offset + c.nf);
    }
    plan->total_out = total_out;
    refresh_meta_geometry(plan);
    build_stage_list(plan);
    return GLS_OK;
}

int gls_graph_update(gls_graph* graph, const gls_plan* plan, void* stream) {
    int rc = require_init();
    if (rc) return rc;
    if (!graph || !plan) return fail(GLS_E_INVALID, "graph or plan is NULL");
    if (graph->plan_id != plan->id)
        return fail(GLS_E_INVALID, "the graph was captured from another plan");
    if (plan->meta.size() != graph->meta_bytes)
        return fail(GLS_E_INVALID, "the plan's descriptors changed size since capture");
    GLS_HIP_TRY(upload_async(graph->workspace, plan->meta.data(), plan->meta.size(),
                             (hipStream_t)stream));
    return GLS_OK;
}

int gls_graph_launch(gls_graph* graph, void* stream) {
    if (!graph || !graph->exec) return fail(GLS_E_INVALID, "graph is NULL");
    GLS_HIP_TRY(hipGraphLaunch(graph->exec, (hipStream_t)stream));
    return GLS_OK;
}

int gls_graph_destroy(gls_graph* graph) {
    if (!graph) return GLS_OK;
    if (graph->exec) (void)hipGraphExecDestroy(graph->exec);
    if (graph->graph) (void)hipGraphDestroy(graph->graph);
    delete graph;
    return GLS_OK;
}

}  // extern "C"
