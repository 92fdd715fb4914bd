// Launch tuner: candidate launch shapes per kernel family, the cost model's
// feature map, measurement, and the choice (learned model or heuristic).
//
// Replaces runIterativeAutoGPU's GAlib search (include/PSkelStencil.hpp:
// 1865-2076), which maximised an analytic objective without timing anything.
#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "pskel_internal.hpp"

using namespace pskel;

namespace pskel {
pskel_status_t mlp_predict_host(const pskel_mlp_t* m, const float* x, size_t n, float* y);
int mlp_inputs(const pskel_mlp_t* m);
}

namespace {

pskel_mlp_t* g_default_model = nullptr;
std::once_flag g_model_once;
// Inference stages inputs in the model's own device buffers and stream, so
// concurrent choices (runIterativeAutoGPU from several host threads) take
// turns on any model.
std::mutex g_predict_mutex;

// The shipped cost model lives next to the library: <libdir>/data/costmodel.bin.
// PSKEL_TUNER_MODEL names another model file, or "none" for no model (the
// tuner then searches by measurement).
void load_shipped_model()
{
    std::string path;
    if (const char* env = std::getenv("PSKEL_TUNER_MODEL")) {
        if (std::strcmp(env, "none") == 0 || env[0] == '\0') return;
        path = env;
    } else {
        Dl_info info;
        if (!dladdr((void*)&load_shipped_model, &info) || !info.dli_fname) return;
        std::string dir(info.dli_fname);
        const size_t slash = dir.rfind('/');
        dir = slash == std::string::npos ? std::string(".") : dir.substr(0, slash);
        path = dir + "/data/costmodel.bin";
    }
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return;
    std::fclose(f);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return;
    pskel_mlp_t* m = nullptr;
    if (pskel_mlp_load(path.c_str(), dev, &m) != PSKEL_OK) return;
    // a model trained on another feature set is not used (the heuristic is)
    if (mlp_inputs(m) == PSKEL_TUNER_NFEAT) g_default_model = m;
    else pskel_mlp_destroy(m);
}

pskel_launch_cfg_t tuned(int tx, int ty, int tz, int k)
{
    pskel_launch_cfg_t c{};
    c.variant = PSKEL_VARIANT_TUNED;
    c.tile_x = tx;
    c.tile_y = ty;
    c.tile_z = tz;
    c.time_tile = k;
    return c;
}

// fused 3D sweeps per launch (depth 2 or 3): tile (cells per lane, rows per
// wave), block (threads, waves along d); the segment length is chosen at launch
pskel_launch_cfg_t fused(int vd, int rw, int threads, int nwd, int depth = 2)
{
    pskel_launch_cfg_t c = tuned(vd, rw, 0, depth);
    c.block_x = threads;
    c.block_y = nwd;
    return c;
}

// the grid a 2D query describes fits the register march (its row pitch is a
// whole number of 1 KiB, as pskel_grid_create lays it out)
bool march2d_fits_query(const pskel_tune_query_t& q)
{
    const size_t pitch = (q.width * 4 + 1023) / 1024 * 1024 / 4;
    return q.ndim == 2 && march2d_ok((int)q.width, (int)q.height, pitch);
}

pskel_launch_cfg_t plain(int bx, int by)
{
    pskel_launch_cfg_t c{};
    c.variant = PSKEL_VARIANT_PLAIN;
    c.block_x = bx;
    c.block_y = by;
    return c;
}

std::vector<pskel_launch_cfg_t> candidates(const pskel_tune_query_t& q, const pskel_stencil_desc_t& d)
{
    std::vector<pskel_launch_cfg_t> v;
    const int pat = classify(d, q.ndim);
    if (q.ndim == 2 && (pat == PAT_J5 || pat == PAT_BOX9 || pat == PAT_LIFE8 || pat == PAT_CLOUD)) {
        const int ry_nwy[][2] = {{1, 2}, {1, 4}, {2, 2}, {2, 4}, {4, 1}, {4, 2}, {4, 4}, {8, 1}, {8, 2}};
        for (auto& t : ry_nwy) v.push_back(tuned(0, t[0], t[1], 1));
        if (pat == PAT_LIFE8) {  // bit-packed tiles: rows per lane x generations per launch
            for (int k = 2; k <= 8; ++k)
                for (int rows : {24, 32, 40, 48}) v.push_back(tuned(0, rows, 0, k));
        } else if (pat != PAT_CLOUD) {  // field stencils run one sweep per launch
            for (int k = 2; k <= 4; ++k)
                for (int cy : {8, 16, 32, 64}) v.push_back(tuned(0, cy, 0, k));
            // the register march (mask range 1): rows per segment (0: about 3 waves per SIMD)
            if (march2d_fits_query(q) && desc_radius(d, 2) == 1)
                for (int k = 2; k <= 4; ++k)
                    for (int seg : {0, 8, 24, 64}) v.push_back(tuned(4, 0, seg, k));
        }
        v.push_back(plain(64, 4));
    } else if (q.ndim == 3 && pat == PAT_J7) {
        const bool narrow = q.storage != PSKEL_DTYPE_F32;
        const int vds16[] = {2, 4, 8}, vds32[] = {1, 2, 4};
        for (int vd : (narrow ? vds16 : vds32))
            for (int rw : {1, 2, 4})
                for (int hs : {16, 32, 64, 128, 256}) {
                    const int lds = 2 * (4 * rw + 2) * (64 * vd + 2 * vd) * 4;
                    if (lds > 64 * 1024) continue;
                    v.push_back(tuned(vd, rw, hs, 1));
                }
        if (q.iterations >= 2) {  // two sweeps per launch: the shapes built for each storage
            v.push_back(fused(4, 3, 768, 2));
            v.push_back(fused(4, 2, 1024, 2));
            v.push_back(fused(4, 3, 768, 1));  // 256-cell rows, 36-row tiles
            v.push_back(fused(2, 3, 768, 1));  // 128-cell rows
            v.push_back(fused(4, 2, 768, 2));  // 12-row tiles of 2 x 6 waves of 2 rows
            if (!narrow) {
                v.push_back(fused(4, 2, 512, 2));
                v.push_back(fused(2, 4, 1024, 4));
            }
        }
        if (q.iterations >= 3 && !narrow) {  // three sweeps per launch (f32)
            v.push_back(fused(4, 4, 512, 2, 3));
            v.push_back(fused(4, 2, 768, 2, 3));
        }
        v.push_back(plain(64, 4));
    } else {
        const int blocks[][2] = {{64, 1}, {64, 2}, {64, 4}, {128, 2}, {256, 1}, {32, 4}};
        for (auto& b : blocks) v.push_back(plain(b[0], b[1]));
    }
    return v;
}

float lg2(double x) { return (float)std::log2(std::max(x, 1.0)); }

// The fraction of a 3D candidate's tile lanes (along d) and rows (along w) that
// hold grid cells: a tile much wider than the grid leaves lanes idle.  1 in 2D.
float fill3d(const pskel_tune_query_t& q, const pskel_launch_cfg_t& c)
{
    if (q.ndim != 3) return 1.f;
    const double D = (double)q.depth, W = (double)q.width;
    auto fit = [](double n, double t) { return n / (std::ceil(n / t) * t); };
    if (c.variant != PSKEL_VARIANT_TUNED)  // plain: block_x along d, block_y along w
        return (float)(fit(D, std::max(c.block_x, 1)) * fit(W, std::max(c.block_y, 1)));
    if (c.time_tile >= 2) {  // fused sweeps: extended tiles overlapping by a lane vector and 2(k-1) rows
        const int vd = c.tile_x > 0 ? c.tile_x : 4, rw = c.tile_y > 0 ? c.tile_y : 3;
        const int nw = c.block_x > 0 ? c.block_x / 64 : 12;
        const int nwd = c.block_y > 0 ? c.block_y % 16 : 2;
        if (nwd <= 0 || nw < nwd) return 1.f;
        const double td = 64.0 * nwd * vd, tw = (double)(nw / nwd) * rw;
        const double ntd = D <= td ? 1.0 : 1.0 + std::ceil((D - td) / (td - 2.0 * vd));
        const double ntw = W <= tw ? 1.0 : 1.0 + std::ceil((W - tw) / (tw - 2.0 * (c.time_tile - 1)));
        return (float)(D / (ntd * td) * W / (ntw * tw));
    }
    // one sweep per launch: tiles of 64 lanes x VD cells by 4 waves x RW rows
    const int vd = c.tile_x > 0 ? c.tile_x : (q.storage == PSKEL_DTYPE_F32 ? 1 : 2);
    const int rw = c.tile_y > 0 ? c.tile_y : 4;
    return (float)(fit(D, 64.0 * vd) * fit(W, 4.0 * rw));
}

void features(const pskel_tune_query_t& q, const pskel_stencil_desc_t& d, const pskel_launch_cfg_t& c, float* f)
{
    std::memset(f, 0, sizeof(float) * PSKEL_TUNER_NFEAT);
    const int pat = classify(d, q.ndim);
    int i = 0;
    f[i++] = lg2((double)q.width);
    f[i++] = lg2((double)q.height);
    f[i++] = lg2((double)(q.ndim == 3 ? q.depth : 1));
    f[i++] = lg2((double)q.width * q.height * (q.ndim == 3 ? q.depth : 1));
    f[i++] = (float)desc_radius(d, q.ndim);
    f[i++] = q.storage == PSKEL_DTYPE_U8 ? 1.f : 0.f;
    f[i++] = q.storage == PSKEL_DTYPE_F32 ? 1.f : 0.f;
    f[i++] = (q.storage == PSKEL_DTYPE_BF16 || q.storage == PSKEL_DTYPE_F16) ? 1.f : 0.f;
    f[i++] = lg2((double)q.iterations);
    f[i++] = pat == PAT_J5 ? 1.f : 0.f;
    f[i++] = pat == PAT_BOX9 ? 1.f : 0.f;
    f[i++] = pat == PAT_LIFE8 ? 1.f : 0.f;
    f[i++] = pat == PAT_J7 ? 1.f : 0.f;
    f[i++] = (float)d.npoints;
    // candidate
    f[i++] = c.variant == PSKEL_VARIANT_TUNED ? 1.f : 0.f;
    f[i++] = lg2((double)std::max(c.tile_x, 1));
    f[i++] = lg2((double)std::max(c.tile_y, 1));
    f[i++] = lg2((double)std::max(c.tile_z, 1));
    f[i++] = (float)std::max(c.time_tile, 1);
    f[i++] = lg2((double)std::max(c.block_x, 1));
    f[i++] = lg2((double)std::max(c.block_y, 1));
    // interactions the model would otherwise have to discover
    f[i++] = (float)std::max(c.time_tile, 1) * (q.ndim == 2 ? 1.f : 0.f) / (float)std::max(c.tile_y, 1);
    f[i++] = lg2((double)q.width * q.height * (q.ndim == 3 ? q.depth : 1) /
                 (double)std::max(1, c.tile_z > 0 && q.ndim == 3 ? c.tile_z : 1));
    f[i++] = fill3d(q, c);
    // fused 16-bit pairs are bound by instruction issue, f32 ones by HBM
    f[i++] = (q.ndim == 3 && c.time_tile >= 2 && q.storage != PSKEL_DTYPE_F32) ? 1.f : 0.f;
    f[i++] = 1.f;  // bias-like constant
}

// Heuristic: the shapes the sweeps on MI355X found best for the large grids.
pskel_launch_cfg_t heuristic(const pskel_tune_query_t& q, const pskel_stencil_desc_t& d)
{
    const int pat = classify(d, q.ndim);
    if (q.ndim == 2 && pat == PAT_CLOUD) return tuned(0, 2, 4, 1);
    if (q.ndim == 2 && pat == PAT_LIFE8) {
        const double cells = (double)q.width * q.height;
        // bit-packed tiles are 1984 columns wide: narrower grids leave lanes idle
        if (q.iterations >= 8 && cells >= (1 << 20) && q.width >= 2048) return tuned(0, 24, 0, 8);
        // small grids: one cell per lane keeps enough blocks in flight
        if (cells < (1 << 18)) return plain(64, 4);
        return tuned(0, 1, 4, 1);
    }
    if (q.ndim == 2 && (pat == PAT_J5 || pat == PAT_BOX9)) {
        // From the tuner's timing table: the register march (4 sweeps,
        // automatic segments) from 0.9M cells for the 5-point stencil and 4M
        // for the 9-point one, which the LDS tiles hold out to longer (they
        // issue its 17 products and sums with less overhead per cell); below,
        // LDS tiles of 16 rows from 0.8M cells and of 8 rows under that, all
        // at 4 sweeps (or the sweep count when smaller)
        const double cells = (double)q.width * q.height;
        if (q.iterations < 2) return tuned(0, 2, 4, 1);
        const int k = (int)std::min<size_t>(q.iterations, 4);
        const double march_from = pat == PAT_J5 ? 0.9e6 : 4.0e6;
        if (cells >= march_from && march2d_fits_query(q) && desc_radius(d, 2) == 1) return tuned(4, 0, 0, k);
        return tuned(0, cells >= 0.8e6 ? 16 : 8, 0, k);
    }
    if (q.ndim == 3 && pat == PAT_J7) {
        // the fused pair whose tiles the grid fills best (512-, 256- or
        // 128-cell rows; the widest on ties); it won on every grid measured,
        // down to 42 deep (29% of a 128-cell row)
        const pskel_launch_cfg_t pairs[] = {fused(4, 3, 768, 2), fused(4, 3, 768, 1), fused(2, 3, 768, 1)};
        int bi = 0;
        float bf = fill3d(q, pairs[0]);
        for (int i = 1; i < 3; ++i) {
            const float f = fill3d(q, pairs[i]);
            if (f > bf + 0.05f) {
                bf = f;
                bi = i;
            }
        }
        // f32 with three sweeps per launch when the grid fills at least 70%
        // of its 512-cell rows (it beat every pair from 406 deep to 512 on the
        // grids measured; at 277 .. 342 deep it lost by up to 5%)
        const double ntd3 = q.depth <= 512 ? 1.0 : 1.0 + std::ceil((q.depth - 512.0) / 504.0);
        if (q.storage == PSKEL_DTYPE_F32 && q.iterations >= 3 && (double)q.depth / (512.0 * ntd3) >= 0.7)
            return fused(4, 4, 512, 2, 3);
        if (q.iterations >= 2 && bf >= 0.2f) return pairs[bi];
        // one sweep per launch: a 64-lane row as wide as the depth needs, and
        // segments short enough for about 1024 blocks
        const bool f32 = q.storage == PSKEL_DTYPE_F32;
        const int vd = f32 ? (q.depth <= 64 ? 1 : q.depth <= 128 ? 2 : 4) : (q.depth <= 128 ? 2 : 4);
        const int rw = f32 ? 1 : 2;
        const double tiles = std::ceil((double)q.depth / (64.0 * vd)) * std::ceil((double)q.width / (4.0 * rw));
        int hs = 16;
        for (int h : {128, 64, 32})
            if (tiles * std::ceil((double)q.height / h) >= 1024) {
                hs = h;
                break;
            }
        return tuned(vd, rw, hs, 1);
    }
    return plain(64, 4);
}

// ---------------------------------------------------------------------------
// Measured search: the fallback without a cost model and the yardstick of its
// regret.  Times are device time per sweep (pskel_tuner_measure); a candidate
// the grid cannot run (a launch error) costs +inf.
// ---------------------------------------------------------------------------
struct Timer {
    pskel_grid_t* g;
    const pskel_stencil_desc_t* d;
    size_t sweeps;
    int reps;
    const std::vector<pskel_launch_cfg_t>* cands;
    std::vector<double> t;  // NaN: not timed yet
    int timed = 0;
    double operator()(size_t i)
    {
        if (std::isnan(t[i])) {
            double ms = 0.0;
            const pskel_status_t st = pskel_tuner_measure(g, d, sweeps, &(*cands)[i], reps, &ms);
            t[i] = st == PSKEL_OK ? ms : std::numeric_limits<double>::infinity();
            ++timed;
        }
        return t[i];
    }
};

size_t search_exhaustive(Timer& time)
{
    size_t best = 0;
    for (size_t i = 0; i < time.t.size(); ++i)
        if (time(i) < time(best)) best = i;
    return best;
}

// A simple GA over candidate indices, the set-up of the reference's GAlib
// search (include/PSkelStencil.hpp:1910-1940: simple GA, binary genome,
// sigma-truncation scaling, one-point crossover, bit-flip mutation) but scoring
// measured time and with a fixed seed (the reference seeds with time(NULL)).
size_t search_genetic(Timer& time, uint64_t seed)
{
    const size_t n = time.t.size();
    int nbits = 1;
    while ((size_t(1) << nbits) < n) ++nbits;
    constexpr int kPop = 16, kGen = 8;
    constexpr double kCross = 0.6, kMut = 0.01;
    uint64_t rs = seed * 0x9E3779B97F4A7C15ull + 1;
    auto rnd = [&rs]() {  // splitmix64 -> [0, 1)
        uint64_t z = (rs += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return (double)((z ^ (z >> 31)) >> 11) * (1.0 / 9007199254740992.0);
    };
    auto decode = [&](uint32_t g) { return (size_t)g % n; };
    std::vector<uint32_t> pop(kPop);
    for (auto& g : pop) {
        g = 0;
        for (int b = 0; b < nbits; ++b) g |= (rnd() < 0.5 ? 1u : 0u) << b;
    }
    size_t best = decode(pop[0]);
    for (auto g : pop)
        if (time(decode(g)) < time(best)) best = decode(g);
    for (int gen = 0; gen < kGen; ++gen) {
        // sigma truncation of the score -time: f = max(score - (mean - 2 sd), 0)
        std::vector<double> score(kPop);
        double mean = 0.0;
        for (int i = 0; i < kPop; ++i) {
            const double t = time(decode(pop[i]));
            score[i] = std::isfinite(t) ? -t : -1e30;
            mean += score[i];
        }
        mean /= kPop;
        double var = 0.0;
        for (double x : score) var += (x - mean) * (x - mean);
        const double sd = std::sqrt(var / kPop);
        std::vector<double> cum(kPop);
        double acc = 0.0;
        for (int i = 0; i < kPop; ++i) {
            acc += std::max(score[i] - (mean - 2.0 * sd), 0.0) + 1e-12;
            cum[i] = acc;
        }
        auto pick = [&]() {
            const double x = rnd() * acc;
            int i = 0;
            while (i < kPop - 1 && cum[i] < x) ++i;
            return pop[i];
        };
        std::vector<uint32_t> next;
        while ((int)next.size() < kPop) {
            uint32_t a = pick(), b = pick();
            if (rnd() < kCross && nbits > 1) {
                const int cut = 1 + (int)(rnd() * (nbits - 1));
                const uint32_t lo = (1u << cut) - 1u;
                const uint32_t na = (a & lo) | (b & ~lo), nb = (b & lo) | (a & ~lo);
                a = na;
                b = nb;
            }
            for (int bit = 0; bit < nbits; ++bit) {
                if (rnd() < kMut) a ^= 1u << bit;
                if (rnd() < kMut) b ^= 1u << bit;
            }
            next.push_back(a & ((1u << nbits) - 1u));
            next.push_back(b & ((1u << nbits) - 1u));
        }
        pop.assign(next.begin(), next.begin() + kPop);
        for (auto g : pop)
            if (time(decode(g)) < time(best)) best = decode(g);
    }
    return best;
}

// Search results of this process, so a shape is searched once.
struct SearchKey {
    std::vector<unsigned char> bytes;
    bool operator<(const SearchKey& o) const { return bytes < o.bytes; }
};
struct SearchResult {
    pskel_launch_cfg_t cfg;
    double ms;
    int32_t timed;
};
std::mutex g_search_mutex;
std::map<SearchKey, SearchResult> g_search_cache;

pskel_status_t check_query(const pskel_tune_query_t* q, const pskel_stencil_desc_t* d)
{
    if (!q || !d) return fail(PSKEL_ERR_INVALID_ARGUMENT, "NULL argument");
    if (q->ndim != 2 && q->ndim != 3) return fail(PSKEL_ERR_INVALID_ARGUMENT, "ndim must be 2 or 3");
    if (d->npoints < 1 || d->npoints > PSKEL_MAX_POINTS) return fail(PSKEL_ERR_INVALID_ARGUMENT, "npoints");
    return PSKEL_OK;
}

}  // namespace

extern "C" {

pskel_status_t pskel_tuner_candidates(const pskel_tune_query_t* q, const pskel_stencil_desc_t* d,
                                      pskel_launch_cfg_t* out, size_t cap, size_t* n)
{
    pskel_status_t st = check_query(q, d);
    if (st != PSKEL_OK) return st;
    const auto v = candidates(*q, *d);
    if (n) *n = v.size();
    if (out)
        for (size_t i = 0; i < v.size() && i < cap; ++i) out[i] = v[i];
    return PSKEL_OK;
}

pskel_status_t pskel_tuner_features(const pskel_tune_query_t* q, const pskel_stencil_desc_t* d,
                                    const pskel_launch_cfg_t* c, float* feat)
{
    pskel_status_t st = check_query(q, d);
    if (st != PSKEL_OK) return st;
    if (!c || !feat) return fail(PSKEL_ERR_INVALID_ARGUMENT, "NULL argument");
    features(*q, *d, *c, feat);
    return PSKEL_OK;
}

pskel_status_t pskel_tuner_heuristic(const pskel_tune_query_t* q, const pskel_stencil_desc_t* d,
                                     pskel_launch_cfg_t* out)
{
    pskel_status_t st = check_query(q, d);
    if (st != PSKEL_OK) return st;
    if (!out) return fail(PSKEL_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = heuristic(*q, *d);
    return PSKEL_OK;
}

pskel_status_t pskel_tuner_set_default_model(pskel_mlp_t* model)
{
    std::call_once(g_model_once, [] {});  // an explicit choice wins over the shipped file
    std::lock_guard<std::mutex> lock(g_predict_mutex);
    g_default_model = model;
    return PSKEL_OK;
}

pskel_status_t pskel_tuner_choose(const pskel_mlp_t* model, const pskel_tune_query_t* q,
                                  const pskel_stencil_desc_t* d, pskel_launch_cfg_t* out)
{
    pskel_status_t st = check_query(q, d);
    if (st != PSKEL_OK) return st;
    if (!out) return fail(PSKEL_ERR_INVALID_ARGUMENT, "out is NULL");
    if (!model) {
        std::call_once(g_model_once, load_shipped_model);
        std::lock_guard<std::mutex> lock(g_predict_mutex);
        model = g_default_model;
    }
    if (!model) {
        *out = heuristic(*q, *d);
        return PSKEL_OK;
    }
    if (mlp_inputs(model) != PSKEL_TUNER_NFEAT)
        return fail(PSKEL_ERR_INVALID_ARGUMENT, "the model's input width is not the tuner's feature count");
    const auto v = candidates(*q, *d);
    std::vector<float> x(v.size() * PSKEL_TUNER_NFEAT), y(v.size());
    for (size_t i = 0; i < v.size(); ++i) features(*q, *d, v[i], &x[i * PSKEL_TUNER_NFEAT]);
    {
        std::lock_guard<std::mutex> lock(g_predict_mutex);
        st = mlp_predict_host(model, x.data(), v.size(), y.data());
    }
    if (st != PSKEL_OK) return st;
    const size_t best = (size_t)(std::min_element(y.begin(), y.end()) - y.begin());
    *out = v[best];
    return PSKEL_OK;
}

pskel_status_t pskel_tuner_has_model(int32_t* out)
{
    if (!out) return fail(PSKEL_ERR_INVALID_ARGUMENT, "out is NULL");
    std::call_once(g_model_once, load_shipped_model);
    std::lock_guard<std::mutex> lock(g_predict_mutex);
    *out = g_default_model ? 1 : 0;
    return PSKEL_OK;
}

pskel_status_t pskel_tuner_search(pskel_grid_t* g, const pskel_stencil_desc_t* d, size_t iterations, int32_t method,
                                  size_t sweeps, int32_t reps, pskel_launch_cfg_t* out, double* ms_per_sweep,
                                  int32_t* measured)
{
    if (!g || !d || !out) return fail(PSKEL_ERR_INVALID_ARGUMENT, "NULL argument");
    if (method != PSKEL_SEARCH_EXHAUSTIVE && method != PSKEL_SEARCH_GENETIC)
        return fail(PSKEL_ERR_INVALID_ARGUMENT, "unknown search method");
    pskel_grid_info_t gi{};
    pskel_status_t st = pskel_grid_info(g, &gi);
    if (st != PSKEL_OK) return st;
    pskel_tune_query_t q{};
    q.ndim = gi.ndim;
    q.storage = gi.storage;
    q.width = gi.width;
    q.height = gi.height;
    q.depth = gi.depth;
    q.iterations = std::max<size_t>(iterations, 1);
    if ((st = check_query(&q, d)) != PSKEL_OK) return st;
    if (sweeps == 0) sweeps = std::min<size_t>(q.iterations, 24);
    reps = std::max(reps, 1);
    SearchKey key;
    auto put = [&key](const void* p, size_t n) {
        key.bytes.insert(key.bytes.end(), (const unsigned char*)p, (const unsigned char*)p + n);
    };
    put(&q, sizeof(q));
    put(d, sizeof(*d));
    put(&method, sizeof(method));
    put(&sweeps, sizeof(sweeps));
    {
        std::lock_guard<std::mutex> lock(g_search_mutex);
        auto it = g_search_cache.find(key);
        if (it != g_search_cache.end()) {
            *out = it->second.cfg;
            if (ms_per_sweep) *ms_per_sweep = it->second.ms;
            if (measured) *measured = 0;
            return PSKEL_OK;
        }
    }
    const auto v = candidates(q, *d);
    Timer time{g, d, sweeps, reps, &v, std::vector<double>(v.size(), std::nan("")), 0};
    const size_t best = method == PSKEL_SEARCH_EXHAUSTIVE ? search_exhaustive(time) : search_genetic(time, 1);
    if (!std::isfinite(time(best))) return fail(PSKEL_ERR_RUNTIME, "no candidate ran on this grid");
    set_error("");
    *out = v[best];
    if (ms_per_sweep) *ms_per_sweep = time(best);
    if (measured) *measured = time.timed;
    std::lock_guard<std::mutex> lock(g_search_mutex);
    g_search_cache[key] = SearchResult{v[best], time(best), time.timed};
    return PSKEL_OK;
}

pskel_status_t pskel_tuner_measure(pskel_grid_t* g, const pskel_stencil_desc_t* d, size_t sweeps,
                                   const pskel_launch_cfg_t* c, int32_t reps, double* ms_per_sweep)
{
    if (!g || !d || !c || !ms_per_sweep || sweeps == 0) return fail(PSKEL_ERR_INVALID_ARGUMENT, "bad argument");
    reps = std::max(reps, 1);
    pskel_run_stats_t st{};
    pskel_status_t s = pskel_run_iterative(g, d, sweeps, c, &st);  // warm-up
    if (s != PSKEL_OK) return s;
    std::vector<double> t;
    for (int r = 0; r < reps; ++r) {
        s = pskel_run_iterative(g, d, sweeps, c, &st);
        if (s != PSKEL_OK) return s;
        t.push_back(st.elapsed_ms / (double)sweeps);
    }
    std::nth_element(t.begin(), t.begin() + t.size() / 2, t.end());
    *ms_per_sweep = t[t.size() / 2];
    return PSKEL_OK;
}

}  // extern "C"
