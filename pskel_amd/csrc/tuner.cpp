// Launch tuner: candidate launch shapes per kernel family, the cost model's
// feature map, measurement, and the choice (learned model or heuristic).
//
// Replaces runIterativeAutoGPU's GAlib search (include/PSkelStencil.hpp:
// 1865-2076), which maximised an analytic objective without timing anything.
#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <cmath>
#include <cstring>
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
void load_shipped_model()
{
    Dl_info info;
    if (!dladdr((void*)&load_shipped_model, &info) || !info.dli_fname) return;
    std::string dir(info.dli_fname);
    const size_t slash = dir.rfind('/');
    dir = slash == std::string::npos ? std::string(".") : dir.substr(0, slash);
    const std::string path = dir + "/data/costmodel.bin";
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

// two fused 3D sweeps per launch: tile (cells per lane, rows per wave), block
// (threads, waves along d); the segment length is chosen at launch
pskel_launch_cfg_t fused(int vd, int rw, int threads, int nwd)
{
    pskel_launch_cfg_t c = tuned(vd, rw, 0, 2);
    c.block_x = threads;
    c.block_y = nwd;
    return c;
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
            for (int k : {2, 4, 6, 8})
                for (int rows : {32, 40, 48}) v.push_back(tuned(0, rows, 0, k));
        } else if (pat != PAT_CLOUD) {  // field stencils run one sweep per launch
            for (int k = 2; k <= 4; ++k)
                for (int cy : {16, 32, 64}) v.push_back(tuned(0, cy, 0, k));
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
            if (!narrow) {
                v.push_back(fused(4, 2, 512, 2));
                v.push_back(fused(2, 4, 1024, 4));
            }
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
    if (c.time_tile >= 2) {  // fused pair: extended tiles overlapping by a lane vector and two rows
        const int vd = c.tile_x > 0 ? c.tile_x : 4, rw = c.tile_y > 0 ? c.tile_y : 3;
        const int nw = c.block_x > 0 ? c.block_x / 64 : 12;
        const int nwd = c.block_y > 0 ? c.block_y % 16 : 2;
        if (nwd <= 0 || nw < nwd) return 1.f;
        const double td = 64.0 * nwd * vd, tw = (double)(nw / nwd) * rw;
        const double ntd = D <= td ? 1.0 : 1.0 + std::ceil((D - td) / (td - 2.0 * vd));
        const double ntw = W <= tw ? 1.0 : 1.0 + std::ceil((W - tw) / (tw - 2.0));
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
        if (q.iterations >= 8 && cells >= (1 << 20) && q.width >= 2048) return tuned(0, 32, 0, 8);
        // small grids: one cell per lane keeps enough blocks in flight
        if (cells < (1 << 18)) return plain(64, 4);
        return tuned(0, 1, 4, 1);
    }
    if (q.ndim == 2 && (pat == PAT_J5 || pat == PAT_BOX9)) {
        // temporal tiles of 16 rows pay at every grid size measured (64 .. 8192 a side)
        if (q.iterations >= 2) return tuned(0, 16, 0, (int)std::min<size_t>(q.iterations, 4));
        return tuned(0, 2, 4, 1);
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
