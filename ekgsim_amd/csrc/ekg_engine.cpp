// ekg_engine.cpp -- host side of the plain and batched simulation paths:
// heart upload, excitation-sequence driver, ECG layout, batch planning and
// the kernel launch sequence.  All validation happens here, before any
// launch, so that every kernel runs on extents it was sized for.
#include "ekg_engine.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <unordered_map>

namespace ekg {

// Default launch shape of the matrix-core path (measured on MI355X at 128^3 x
// 256 requests: 128-voxel chunks in 2-wave blocks keep eight blocks per CU;
// larger blocks leave waves idle in the transition tiles while holding LDS).
constexpr int DEFAULT_CHUNK = 128;
constexpr int DEFAULT_WAVES = 2;


[[noreturn]] void fail(int code, const std::string& msg) { throw Error{code, msg}; }

void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess) fail(EKG_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

void DevBuf::reserve(size_t n)
{
    if (n <= bytes) return;
    release();
    if (n == 0) return;
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, n);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        fail(EKG_ENOMEM, "device allocation of " + std::to_string(n) + " bytes failed: " +
                             hipGetErrorString(e));
    }
    p = q;
    bytes = n;
}

void DevBuf::adopt(void* q, size_t n)
{
    release();
    p = q;
    bytes = n;
    owned = false;
}

void DevBuf::release()
{
    if (p && owned) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    owned = true;
}

void Workspace::need(DevBuf& b, size_t n)
{
    if (!arena) {
        // grow by at least half the current size: a workspace that serves
        // groups of varying size reallocates (and synchronises the device
        // in hipFree) a few times, not at every larger group
        if (n > b.bytes) b.reserve(std::max(n, b.bytes + b.bytes / 2));
        return;
    }
    const size_t at = (arena_used + 255) & ~(size_t)255;
    if (at > arena_bytes || n > arena_bytes - at)
        fail(EKG_ENOMEM, "the group needs more than the " + std::to_string(arena_bytes >> 20) +
                             " MiB static device workspace of the allocation-free entry point; use fewer "
                             "requests per call or ekg_simulate_batch_device");
    b.adopt(arena + at, n);
    arena_used = at + n;
}

void Ctx::bind() const { hip_check(hipSetDevice(device), "hipSetDevice"); }

Workspace::~Workspace()
{
    if (graph) (void)hipGraphExecDestroy(graph);
    if (side) {
        (void)hipStreamSynchronize(side);
        (void)hipStreamDestroy(side);
    }
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
}

void Workspace::ensure_side()
{
    if (side) return;
    hip_check(hipStreamCreateWithFlags(&side, hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming), "hipEventCreate");
    hip_check(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming), "hipEventCreate");
}

hipEvent_t Ctx::get_event()
{
    if (!free_events.empty()) {
        hipEvent_t e = free_events.back();
        free_events.pop_back();
        return e;
    }
    hipEvent_t e;
    hip_check(hipEventCreate(&e), "hipEventCreate");
    return e;
}

void Ctx::collect_timings()
{
    for (auto& pr : pending) {
        hip_check(hipEventSynchronize(pr.second), "hipEventSynchronize");
        float ms = 0.f;
        hip_check(hipEventElapsedTime(&ms, pr.first, pr.second), "hipEventElapsedTime");
        fused_ms += ms;
        ++fused_launches;
        free_events.push_back(pr.first);
        free_events.push_back(pr.second);
    }
    pending.clear();
}

// Neighbourhood::create + callbacks (simulator.h:325-390): offsets in the
// order of active over {0,1,2}^3 with active[0] outermost.
int nhood_offsets(int type, int8_t dif[26][3])
{
    int n = 0;
    for (int a0 = 0; a0 < 3; ++a0)
        for (int a1 = 0; a1 < 3; ++a1)
            for (int a2 = 0; a2 < 3; ++a2) {
                const int d0 = std::abs(a0 - 1), d1 = std::abs(a1 - 1), d2 = std::abs(a2 - 1);
                const int mn12 = std::min(d1, d2), mx12 = std::max(d1, d2);
                const int mn = std::min(d0, mn12), mx = std::max(d0, mx12);
                bool take = false;
                switch (type) {
                case EKG_NHOOD_2D4: take = mn12 == 1 && mx12 == 1 && d0 == 0; break;
                case EKG_NHOOD_2D8: take = mx12 == 1 && d0 == 0; break;
                case EKG_NHOOD_3D4: take = mn == 1 && mx == 1; break;
                case EKG_NHOOD_3D8: take = mx == 1; break;
                default: return -1;
                }
                if (take) {
                    dif[n][0] = (int8_t)(a0 - 1);
                    dif[n][1] = (int8_t)(a1 - 1);
                    dif[n][2] = (int8_t)(a2 - 1);
                    ++n;
                }
            }
    return n;
}

int64_t num_steps(double length, double dt)
{
    if (!(dt > 0) || !std::isfinite(dt) || !std::isfinite(length) || length <= 0) return 0;
    const double s = std::ceil(length / dt);  // simulator.cpp:470
    if (!(s < 1e9)) return -1;
    return (int64_t)s;
}

// Iterations of the neighbour loop `for (simTime = start; simTime < start +
// length; simTime += dt)` (simulator.cpp:481, 519).  It can differ from T
// for steps like 1/3; samples at or beyond it receive no dipole.
static int64_t neighbour_steps(double start, double length, double dt, int64_t cap)
{
    const double total = length + start;
    int64_t n = 0;
    for (double st = start; st < total; st += dt) {
        if (++n > cap) break;
    }
    return n;
}

ReqSpec make_spec(Heart& h, const ekg_sim_desc& d, bool host_values)
{
    ReqSpec r;
    r.h = &h;
    r.L = d.n_leads;
    r.nl = d.n_layers;
    r.nhood = d.nhood;
    r.start = d.start;
    r.length = d.length;
    r.dt = d.dt;
    if (r.L < 1 || r.L > EKG_MAX_LEADS)
        fail(EKG_EINVAL, "n_leads must be in 1.." + std::to_string(EKG_MAX_LEADS));
    if (r.nhood < EKG_NHOOD_2D4 || r.nhood > EKG_NHOOD_3D8)
        fail(EKG_EINVAL, "unknown neighbourhood (only know of these: 2D4, 3D4, 2D8, 3D8 (cube))");
    if (r.nl < h.max_layer || r.nl < 1 || r.nl > 255)
        fail(EKG_EINVAL, "the heart has " + std::to_string(h.max_layer) + " layers but " +
                             std::to_string(r.nl) + " layer APs were given");
    if (!(r.dt > 0) || !std::isfinite(r.dt) || !std::isfinite(r.start) || !std::isfinite(r.length))
        fail(EKG_EINVAL, "time step must be positive and times finite");
    r.T = num_steps(r.length, r.dt);
    if (r.T < 0 || r.T > (1 << 24)) fail(EKG_EINVAL, "too many time steps");
    r.Tp = std::max<int64_t>(64, ((r.T + 63) / 64) * 64);
    r.n_tiles = (int)(r.Tp / 64);
    r.Tn = neighbour_steps(r.start, r.length, r.dt, r.T + 1);
    if (r.Tn > r.T)
        fail(EKG_EINVAL, "time step makes the neighbour loop of the reference run past the "
                         "sample buffer (simulator.cpp:519); choose another time step");
    if (host_values) {
        if (!d.layer_k || !d.lead_pos) fail(EKG_EINVAL, "layer_k and lead_pos are required");
        for (int i = 0; i < 9 * r.nl; ++i)
            if (!std::isfinite(d.layer_k[i])) fail(EKG_EINVAL, "non-finite AP coefficient");
        for (int i = 0; i < 3 * r.L; ++i)
            if (!std::isfinite(d.lead_pos[i])) fail(EKG_EINVAL, "non-finite lead position");
        r.k = d.layer_k;
        r.pos = d.lead_pos;
    }
    return r;
}

bool heart_mfma(const Ctx& ctx, const Heart& h)
{
    int kernel = h.tune_kernel ? h.tune_kernel : ctx.kernel;
    if (ctx.kernel != EKG_KERNEL_AUTO) kernel = ctx.kernel;  // as in plan_group
    return kernel != EKG_KERNEL_VALU;
}

GroupPlan plan_group(Ctx& ctx, std::vector<ReqSpec> reqs)
{
    if (reqs.empty()) fail(EKG_EINVAL, "empty group");
    GroupPlan g;
    g.nhood = reqs[0].nhood;
    int L = 1, kernel = ctx.kernel, waves = 0, phase = 0;
    const bool path0 = heart_mfma(ctx, *reqs[0].h);
    for (const ReqSpec& r : reqs) {
        if (r.nhood != g.nhood) fail(EKG_EINVAL, "all requests of a launch group must share the neighbourhood");
        // each heart's layout is the one of its own contraction path
        if (heart_mfma(ctx, *r.h) != path0)
            fail(EKG_EINVAL, "all requests of a launch group must take the same contraction path");
        L = std::max(L, r.L);
        if (r.h->tune_kernel) kernel = r.h->tune_kernel;
        if (r.h->tune_waves) waves = std::max(waves, r.h->tune_waves);
        if (r.h->tune_phase && !phase) phase = r.h->tune_phase;
    }
    if (ctx.kernel != EKG_KERNEL_AUTO) kernel = ctx.kernel;  // an explicit context choice wins
    g.LP = L <= 4 ? 4 : L <= 8 ? 8 : L <= 12 ? 12 : 16;
    g.mfma = kernel != EKG_KERNEL_VALU;
    g.waves = waves ? waves : DEFAULT_WAVES;
    g.lf_phase = phase ? phase : 2;
    g.reqs = std::move(reqs);
    return g;
}

// Largest activation-time span inside one chunk (ms).  The saturated AP path
// scales exponentials by e^{(k4+k5)(a - t_ref)} per voxel; a bounded span
// keeps that factor far from FP32 overflow for any realistic k4, k5.
constexpr double CHUNK_SPAN_MS = 16.0;
// Smallest activation class by default: a run of at least this many voxels of
// one layer with one activation time takes the class path.
constexpr int DEFAULT_CLASS_MIN = 64;
// Largest class chunk (voxels): a class is kept as a run of class chunks
// (layout bookkeeping; the class-window kernel takes whole classes).
constexpr int64_t CLASS_CHUNK = 4096;
// Classes per class window and groups of 16 entries per piece (see ensure_cells).
constexpr int CLASS_WINDOW = ekgk::CLASS_WINDOW;
constexpr int CLASS_PIECE = ekgk::CLASS_PIECE;

namespace {
// Chunk size: DEFAULT_CHUNK voxels unless the heart is tuned.  Chunks are
// layer-pure and span at most CHUNK_SPAN_MS of activation time.
int64_t chunk_target_of(const Heart& h) { return h.tune_chunk ? h.tune_chunk : DEFAULT_CHUNK; }
// Activation classes (matrix-core path only): runs of one layer and one
// activation time with at least this many voxels (0: none).
int class_min_of(const Heart& h, bool mfma)
{
    return (!mfma || h.tune_class < 0) ? 0 : (h.tune_class > 0 ? h.tune_class : DEFAULT_CLASS_MIN);
}
}  // namespace

bool layout_ready(const Heart& h, bool mfma, int nhood)
{
    const EcgLayout& L = h.layout;
    if (!(h.excited && L.valid && L.chunk_target == chunk_target_of(h) && L.class_min == class_min_of(h, mfma)))
        return false;
    return !mfma || (nhood >= 0 && nhood < 4 && L.cells_built[nhood]);
}

void ensure_layout(Heart& h, bool mfma)
{
    if (!h.excited) fail(EKG_ESTATE, "excitation sequence not computed or loaded");
    EcgLayout& L = h.layout;
    const int64_t target = chunk_target_of(h);
    const int cmin = class_min_of(h, mfma);
    if (L.valid && L.chunk_target == target && L.class_min == cmin) return;
    // The layout's device buffers are rewritten in place: kernels still
    // reading the old layout on any stream (a server's slots) finish first.
    if (L.valid) {
        h.ctx->bind();
        hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    }
    L.valid = false;
    std::vector<int64_t> idx;
    idx.reserve(h.n_myo);
    for (int64_t i = 0; i < h.nvox; ++i)
        if (h.h_layer[i]) idx.push_back(i);
    std::stable_sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) {
        if (h.h_layer[a] != h.h_layer[b]) return h.h_layer[a] < h.h_layer[b];
        return h.h_act[a] < h.h_act[b];
    });
    const int64_t V = (int64_t)idx.size();
    L.V = V;
    L.max_layer = h.max_layer;
    L.h_chunks.clear();
    L.h_cchunks.clear();
    L.h_classes.clear();
    L.h_class_a.clear();
    L.chunk_target = (int)target;
    L.class_min = cmin;
    for (bool& b : L.cells_built) b = false;
    // reference time of layer l: its earliest activation time, rounded to FP32
    L.h_layer_t0.assign(h.max_layer + 1, 0.0);
    std::vector<int64_t> order;  // the layout's voxel order
    order.reserve(V);
    auto tbits_of = [](double a) {
        const float t = (float)a;
        int b;
        std::memcpy(&b, &t, 4);
        return b;
    };
    int64_t v = 0;
    while (v < V) {
        const int lay = h.h_layer[idx[v]];
        int64_t e = v;
        while (e < V && h.h_layer[idx[e]] == lay) ++e;
        L.h_layer_t0[lay] = (double)(float)h.h_act[idx[v]];
        // voxels outside classes first, then the classes' runs
        std::vector<std::pair<int64_t, int64_t>> runs;
        const int64_t m0 = (int64_t)order.size();
        for (int64_t s0 = v; s0 < e;) {
            int64_t s1 = s0 + 1;
            while (s1 < e && h.h_act[idx[s1]] == h.h_act[idx[s0]]) ++s1;
            if (cmin > 0 && s1 - s0 >= cmin)
                runs.push_back({s0, s1});
            else
                order.insert(order.end(), idx.begin() + s0, idx.begin() + s1);
            s0 = s1;
        }
        const int64_t m1 = (int64_t)order.size();
        // chunks: k near-equal pieces of at most `target`, each further cut
        // where the activation span would exceed CHUNK_SPAN_MS
        const int64_t n = m1 - m0;
        const int64_t k = (n + target - 1) / target;
        for (int64_t c = 0; c < k; ++c) {
            int64_t b0 = m0 + n * c / k;
            const int64_t b1 = m0 + n * (c + 1) / k;
            while (b0 < b1) {
                const float tref = (float)h.h_act[order[b0]];
                int64_t b2 = b0 + 1;
                while (b2 < b1 && h.h_act[order[b2]] - (double)tref <= CHUNK_SPAN_MS) ++b2;
                L.h_chunks.push_back(int4{lay, (int)b0, (int)b2, tbits_of(h.h_act[order[b0]])});
                b0 = b2;
            }
        }
        // class chunks: near-equal pieces of at most CLASS_CHUNK per class
        for (const auto& rr : runs) {
            const int64_t c0 = (int64_t)order.size(), nc = rr.second - rr.first;
            order.insert(order.end(), idx.begin() + rr.first, idx.begin() + rr.second);
            const int64_t kc = (nc + CLASS_CHUNK - 1) / CLASS_CHUNK;
            const int first = (int)L.h_cchunks.size();
            const double a = h.h_act[idx[rr.first]];
            for (int64_t c = 0; c < kc; ++c)
                L.h_cchunks.push_back(int4{lay, (int)(c0 + nc * c / kc), (int)(c0 + nc * (c + 1) / kc), tbits_of(a)});
            L.h_classes.push_back(int4{lay, first, (int)L.h_cchunks.size(), 0});
            L.h_class_a.push_back(a);
        }
        v = e;
    }
    idx.swap(order);
    L.n_chunks = (int)L.h_chunks.size();
    L.n_cchunks = (int)L.h_cchunks.size();
    L.n_classes = (int)L.h_classes.size();
    // chunks of layer l: [layer_chunk[l], layer_chunk[l+1])
    L.h_layer_chunk.assign(h.max_layer + 2, L.n_chunks);
    for (int c = L.n_chunks - 1; c >= 0; --c) L.h_layer_chunk[L.h_chunks[c].x] = c;
    L.h_layer_chunk[0] = 0;
    for (int l = h.max_layer; l >= 1; --l)
        if (L.h_layer_chunk[l] > L.h_layer_chunk[l + 1]) L.h_layer_chunk[l] = L.h_layer_chunk[l + 1];
    // per sorted voxel: chunk index and activation time (FP64 and hi/lo FP32)
    std::vector<int> vchunk(V ? V : 1);
    for (int c = 0; c < L.n_chunks; ++c)
        for (int v2 = L.h_chunks[c].y; v2 < L.h_chunks[c].z; ++v2) vchunk[v2] = c;
    std::vector<double> a64(V ? V : 1);
    std::vector<float2> af(V ? V : 1);
    std::vector<uint2> geo(V ? V : 1);
    if (h.n0 > 1024 || h.n1 > 1024 || h.n2 > 1024) fail(EKG_EINVAL, "grid dimensions above 1024");
    for (int64_t i = 0; i < V; ++i) {
        const double a = h.h_act[idx[i]];
        a64[i] = a;
        const float hi = (float)a;
        af[i] = float2{hi, (float)(a - (double)hi)};
        // grid coordinates and the 3x3x3 myocardium mask of the voxel
        const int64_t x = idx[i];
        const int i2 = (int)(x % h.n2);
        const int i1 = (int)((x / h.n2) % h.n1), i0 = (int)(x / ((int64_t)h.n2 * h.n1));
        uint32_t mask = 0;
        for (int d0 = -1; d0 <= 1; ++d0)
            for (int d1 = -1; d1 <= 1; ++d1)
                for (int d2 = -1; d2 <= 1; ++d2) {
                    const int z = i0 + d0, y = i1 + d1, xx = i2 + d2;
                    if (z < 0 || z >= h.n0 || y < 0 || y >= h.n1 || xx < 0 || xx >= h.n2) continue;
                    if (h.h_layer[((int64_t)z * h.n1 + y) * h.n2 + xx])
                        mask |= 1u << ((d0 + 1) * 9 + (d1 + 1) * 3 + (d2 + 1));
                }
        geo[i] = make_uint2((uint32_t)i0 | ((uint32_t)i1 << 10) | ((uint32_t)i2 << 20), mask);
    }
    // 8^3 bricks of the grid that hold myocardium, with their voxels
    // {sorted index, local cell} grouped by brick (counting sort)
    const int nb0 = (h.n0 + 7) / 8, nb1 = (h.n1 + 7) / 8, nb2 = (h.n2 + 7) / 8;
    std::vector<int> cnt((size_t)nb0 * nb1 * nb2 + 1, 0);
    std::vector<int> vbrick(V ? V : 1);
    for (int64_t i = 0; i < V; ++i) {
        const int64_t x = idx[i];
        const int i2 = (int)(x % h.n2);
        const int64_t r = x / h.n2;
        const int i1 = (int)(r % h.n1), i0 = (int)(r / h.n1);
        vbrick[i] = ((i0 / 8) * nb1 + i1 / 8) * nb2 + i2 / 8;
        ++cnt[vbrick[i]];
    }
    std::vector<int4> bricks;
    std::vector<int> boff;
    std::vector<int> bslot(cnt.size(), -1);
    for (int bq = 0; bq < nb0 * nb1 * nb2; ++bq)
        if (cnt[bq]) {
            bslot[bq] = (int)bricks.size();
            bricks.push_back(int4{(bq / (nb1 * nb2)) * 8, ((bq / nb2) % nb1) * 8, (bq % nb2) * 8, 0});
        }
    boff.assign(bricks.size() + 1, 0);
    for (int bq = 0; bq < nb0 * nb1 * nb2; ++bq)
        if (cnt[bq]) boff[bslot[bq] + 1] = cnt[bq];
    for (size_t q = 1; q < boff.size(); ++q) boff[q] += boff[q - 1];
    std::vector<int2> bvox(V ? V : 1);
    std::vector<int> fill(boff.begin(), boff.end() - 1);
    for (int64_t i = 0; i < V; ++i) {
        const int64_t x = idx[i];
        const int i2 = (int)(x % h.n2);
        const int64_t r = x / h.n2;
        const int i1 = (int)(r % h.n1), i0 = (int)(r / h.n1);
        const int cell = ((i0 % 8) * 8 + (i1 % 8)) * 8 + (i2 % 8);
        bvox[fill[bslot[vbrick[i]]]++] = int2{(int)i, cell};
    }
    // cell order inside each brick: a wave's stencil gathers then touch
    // neighbouring LDS words (few bank conflicts)
    for (size_t q = 0; q + 1 < boff.size(); ++q)
        std::sort(bvox.begin() + boff[q], bvox.begin() + boff[q + 1],
                  [](const int2& x, const int2& y) { return x.y < y.y; });
    L.n_bricks = (int)bricks.size();
    // myocardial cells of each brick's 10^3 box (brick + 1-voxel halo)
    std::vector<int> hoff(bricks.size() + 1, 0);
    std::vector<uint16_t> hcells;
    for (size_t q = 0; q < bricks.size(); ++q) {
        const int4 bq = bricks[q];
        for (int c = 0; c < 1000; ++c) {
            const int gz = bq.x - 1 + c / 100, gy = bq.y - 1 + (c / 10) % 10, gx = bq.z - 1 + c % 10;
            if (gz < 0 || gz >= h.n0 || gy < 0 || gy >= h.n1 || gx < 0 || gx >= h.n2) continue;
            if (h.h_layer[((int64_t)gz * h.n1 + gy) * h.n2 + gx]) hcells.push_back((uint16_t)c);
        }
        hoff[q + 1] = (int)hcells.size();
    }
    if (hcells.empty()) hcells.push_back(0);
    h.ctx->bind();
    L.vox_chunk.reserve(vchunk.size() * sizeof(int));
    L.bricks.reserve(std::max<size_t>(1, bricks.size()) * sizeof(int4));
    L.brick_off.reserve(boff.size() * sizeof(int));
    L.brick_vox.reserve(bvox.size() * sizeof(int2));
    L.vox_a64.reserve(a64.size() * sizeof(double));
    L.vox_a.reserve(af.size() * sizeof(float2));
    L.vox_geo.reserve(geo.size() * sizeof(uint2));
    L.chunks.reserve(std::max<size_t>(1, L.h_chunks.size()) * sizeof(int4));
    hip_check(hipMemcpy(L.vox_chunk.p, vchunk.data(), vchunk.size() * sizeof(int), hipMemcpyHostToDevice),
              "upload voxel chunks");
    if (!bricks.empty())
        hip_check(hipMemcpy(L.bricks.p, bricks.data(), bricks.size() * sizeof(int4), hipMemcpyHostToDevice),
                  "upload bricks");
    hip_check(hipMemcpy(L.brick_off.p, boff.data(), boff.size() * sizeof(int), hipMemcpyHostToDevice),
              "upload brick offsets");
    hip_check(hipMemcpy(L.brick_vox.p, bvox.data(), bvox.size() * sizeof(int2), hipMemcpyHostToDevice),
              "upload brick voxels");
    L.halo_off.reserve(hoff.size() * sizeof(int));
    L.halo_cells.reserve(hcells.size() * sizeof(uint16_t));
    hip_check(hipMemcpy(L.halo_off.p, hoff.data(), hoff.size() * sizeof(int), hipMemcpyHostToDevice),
              "upload halo offsets");
    hip_check(hipMemcpy(L.halo_cells.p, hcells.data(), hcells.size() * sizeof(uint16_t),
                        hipMemcpyHostToDevice),
              "upload halo cells");
    hip_check(hipMemcpy(L.vox_a64.p, a64.data(), a64.size() * sizeof(double), hipMemcpyHostToDevice),
              "upload activation times");
    hip_check(hipMemcpy(L.vox_a.p, af.data(), af.size() * sizeof(float2), hipMemcpyHostToDevice),
              "upload activation times");
    hip_check(hipMemcpy(L.vox_geo.p, geo.data(), geo.size() * sizeof(uint2), hipMemcpyHostToDevice),
              "upload voxel geometry");
    L.h_geo.swap(geo);
    if (!L.h_chunks.empty())
        hip_check(hipMemcpy(L.chunks.p, L.h_chunks.data(), L.h_chunks.size() * sizeof(int4),
                            hipMemcpyHostToDevice),
                  "upload chunks");
    L.cchunks.reserve(std::max<size_t>(1, L.h_cchunks.size()) * sizeof(int4));
    L.classes.reserve(std::max<size_t>(1, L.h_classes.size()) * sizeof(int4));
    L.class_a.reserve(std::max<size_t>(1, L.h_class_a.size()) * sizeof(double));
    if (!L.h_cchunks.empty()) {
        hip_check(hipMemcpy(L.cchunks.p, L.h_cchunks.data(), L.h_cchunks.size() * sizeof(int4),
                            hipMemcpyHostToDevice),
                  "upload class chunks");
        hip_check(hipMemcpy(L.classes.p, L.h_classes.data(), L.h_classes.size() * sizeof(int4),
                            hipMemcpyHostToDevice),
                  "upload classes");
        hip_check(hipMemcpy(L.class_a.p, L.h_class_a.data(), L.h_class_a.size() * sizeof(double),
                            hipMemcpyHostToDevice),
                  "upload class times");
    }
    L.layer_chunk.reserve(L.h_layer_chunk.size() * sizeof(int));
    L.layer_t0.reserve(L.h_layer_t0.size() * sizeof(double));
    hip_check(hipMemcpy(L.layer_chunk.p, L.h_layer_chunk.data(), L.h_layer_chunk.size() * sizeof(int),
                        hipMemcpyHostToDevice),
              "upload layer chunks");
    hip_check(hipMemcpy(L.layer_t0.p, L.h_layer_t0.data(), L.h_layer_t0.size() * sizeof(double),
                        hipMemcpyHostToDevice),
              "upload layer times");
    ekgk::HeartDev& d = L.dev;
    d.grid_layer = h.d_layer.as<uint8_t>();
    d.bricks = L.bricks.as<int4>();
    d.brick_off = L.brick_off.as<int>();
    d.brick_vox = L.brick_vox.as<int2>();
    d.halo_off = L.halo_off.as<int>();
    d.halo_cells = L.halo_cells.as<uint16_t>();
    d.vox_chunk = L.vox_chunk.as<int>();
    d.vox_a64 = L.vox_a64.as<double>();
    d.vox_a = L.vox_a.as<float2>();
    d.vox_geo = L.vox_geo.as<uint2>();
    d.chunks = L.chunks.as<int4>();
    d.layer_chunk = L.layer_chunk.as<int>();
    d.layer_t0 = L.layer_t0.as<double>();
    d.cchunks = L.cchunks.as<int4>();
    d.classes = L.classes.as<int4>();
    d.class_a = L.class_a.as<double>();
    d.n0 = h.n0;
    d.n1 = h.n1;
    d.n2 = h.n2;
    d.V = (int)V;
    d.n_chunks = L.n_chunks;
    d.n_bricks = L.n_bricks;
    d.max_layer = h.max_layer;
    d.n_cchunks = L.n_cchunks;
    d.n_classes = L.n_classes;
    L.valid = true;
}

// The sum of the G rows of a set of voxels S, regrouped by cell: a voxel u
// contributes m(u) . s_u (s_u = sum of the d in N with u + d myocardial) and
// m(u + d) . d for each such d, so
//   sum_{u in S} G(u) = sum_w m(w) . D(w),
//   D(w) = [w in S] s_w + sum_{d in N, w - d in S, w myocardial} d,
// an integer vector per cell; cells inside S cancel (D = 0).  With S an
// activation class, every voxel of S has the same AP, so the class needs only
// this sum; a cell w has grid coordinates + 1 (the reference's voxel centre)
// c(w), D(w) and its moment K(w) = -D(w) . c(w), all integers.
//
// A cell borders several classes (4 to 5 on average on a heart with delays of
// 1 to 3 ms per voxel), and its lead weight s(w) = |p - c(w)|^-3 is the same
// for all of them.  So the classes are taken CLASS_WINDOW at a time (in layer
// and time order): an entry is a cell with a non-zero D for at least one
// class of the window, and the class-window kernel evaluates each entry's
// weight once per lead and request and adds it, times the entry's D and K for
// class i of the window, to that class's sums.  Entries are sorted by their
// class mask and cut into groups of 16; a group lists, for every class of its
// mask (the union of its entries'), the 64 lanes' coefficients, and the
// kernel runs one matrix product per class of the mask.
//
// Each of the 64 lanes feeds the matrix products fixed operands of its group
// (lane = 16 kq + 4 b + rc, entry index within the group el = 4 b + kq,
// ea = 4 b + rc):
//   av = coordinate kq of entry ea (kq = 3: 1),
//   cc = |c|^2 of entry el,
//   bv = coefficient rc of entry el for class i: D_0, D_1, D_2 (rc < 3) or K.
// They are stored lane-major, as FP32 (every value is an integer below 2^24,
// exact): group q occupies rows crow[q].x, ... of 64 floats in cgrp: av, cc,
// then bv for each class of its mask in order; crow[q].y is the mask.  A
// group's missing entries (past the window's last) repeat its first entry
// with coefficients 0, so their weights are finite and add nothing.
//
// For parallelism a window's groups are cut into the fewest pieces of at
// most CLASS_PIECE groups, of equal size up to one group (so that no block
// is left with a short remainder), one block per (piece, request); each writes its sums
// for all CLASS_WINDOW classes, and class_sums_kernel adds a class's pieces
// in order.  The cut depends on the heart only, so a request's ECG does not
// depend on the batch it runs in.
void ensure_cells(Heart& h, int nhood)
{
    EcgLayout& L = h.layout;
    if (nhood < 0 || nhood > 3) fail(EKG_EINVAL, "unknown neighbourhood");
    if (L.cells_built[nhood]) return;
    int8_t dif[26][3];
    const int nd = nhood_offsets(nhood, dif);
    // cell -> entry of the current window (-1: none), reset after each
    // window: a dense array over the grid, or a hash map for grids above
    // 2^27 cells (the array would exceed 512 MB)
    const bool dense = h.nvox <= ((int64_t)1 << 27);
    std::vector<int> slot(dense ? (size_t)h.nvox : 0, -1);
    std::unordered_map<int64_t, int> slot_map;
    std::vector<int64_t> touched;   // cells of the current class, first-touch order
    std::vector<int> acc;           // D per touched cell, three ints
    std::vector<int> tslot(dense ? (size_t)h.nvox : 0, -1);  // cell -> touched index of the current class
    std::unordered_map<int64_t, int> tslot_map;
    struct Entry {
        int c[3];
        int mask;
        int first;  // first touch (order within a mask)
        int coef[CLASS_WINDOW][4];
    };
    std::vector<Entry> ent;
    std::vector<float> grp;          // rows of 64 floats
    std::vector<int2> crow;          // per group: first row, class mask
    std::vector<int4> pieces;        // {window, first group, end group, 0}
    std::vector<int> wpiece(1, 0);   // pieces of window w: [wpiece[w], wpiece[w+1])
    int64_t n_cells = 0, n_entries = 0;
    auto bit = [](int d0, int d1, int d2) { return (d0 + 1) * 9 + (d1 + 1) * 3 + (d2 + 1); };
    auto add = [&](int z, int y, int x, int d0, int d1, int d2) {
        const int64_t gi = ((int64_t)z * h.n1 + y) * h.n2 + x;
        int sl;
        if (dense) {
            sl = tslot[gi];
            if (sl < 0) sl = tslot[gi] = (int)touched.size();
        } else {
            sl = tslot_map.emplace(gi, (int)touched.size()).first->second;
        }
        if (sl == (int)touched.size()) {
            touched.push_back(gi);
            acc.insert(acc.end(), {0, 0, 0});
        }
        acc[3 * sl] += d0;
        acc[3 * sl + 1] += d1;
        acc[3 * sl + 2] += d2;
    };
    const int n_win = (L.n_classes + CLASS_WINDOW - 1) / CLASS_WINDOW;
    for (int w = 0; w < n_win; ++w) {
        ent.clear();
        const int k0 = w * CLASS_WINDOW, k1 = std::min(L.n_classes, k0 + CLASS_WINDOW);
        for (int k = k0; k < k1; ++k) {
            // the class's voxels: those of its class chunks, which are consecutive
            const int4 cl = L.h_classes[k];
            const int v_begin = L.h_cchunks[cl.y].y, v_end = L.h_cchunks[cl.z - 1].z;
            touched.clear();
            acc.clear();
            for (int v = v_begin; v < v_end; ++v) {
                const uint2 g = L.h_geo[v];
                const int z = (int)(g.x & 1023u), y = (int)((g.x >> 10) & 1023u), x = (int)(g.x >> 20);
                int s0 = 0, s1 = 0, s2 = 0;
                for (int q = 0; q < nd; ++q) {
                    if (!((g.y >> bit(dif[q][0], dif[q][1], dif[q][2])) & 1u)) continue;
                    add(z + dif[q][0], y + dif[q][1], x + dif[q][2], dif[q][0], dif[q][1], dif[q][2]);
                    s0 += dif[q][0];
                    s1 += dif[q][1];
                    s2 += dif[q][2];
                }
                add(z, y, x, s0, s1, s2);
            }
            tslot_map.clear();
            for (size_t t = 0; t < touched.size(); ++t) {
                const int64_t gi = touched[t];
                if (dense) tslot[gi] = -1;
                const int D0 = acc[3 * t], D1 = acc[3 * t + 1], D2 = acc[3 * t + 2];
                if (!(D0 || D1 || D2)) continue;
                if (std::abs(D0) > 127 || std::abs(D1) > 127 || std::abs(D2) > 127)
                    fail(EKG_EINVAL, "dipole vector out of range");
                ++n_cells;
                int e;
                if (dense) {
                    e = slot[gi];
                    if (e < 0) e = slot[gi] = (int)ent.size();
                } else {
                    e = slot_map.emplace(gi, (int)ent.size()).first->second;
                }
                if (e == (int)ent.size()) {
                    const int x = (int)(gi % h.n2), y = (int)((gi / h.n2) % h.n1),
                              z = (int)(gi / ((int64_t)h.n2 * h.n1));
                    Entry en{};
                    en.c[0] = z + 1;
                    en.c[1] = y + 1;
                    en.c[2] = x + 1;
                    en.first = e;
                    ent.push_back(en);
                }
                Entry& en = ent[e];
                const int i = k - k0;
                en.mask |= 1 << i;
                en.coef[i][0] = D0;
                en.coef[i][1] = D1;
                en.coef[i][2] = D2;
                en.coef[i][3] = -(D0 * en.c[0] + D1 * en.c[1] + D2 * en.c[2]);
            }
        }
        // reset the cell -> entry map for the next window
        if (dense)
            for (const Entry& en : ent)
                slot[((int64_t)(en.c[0] - 1) * h.n1 + (en.c[1] - 1)) * h.n2 + (en.c[2] - 1)] = -1;
        else
            slot_map.clear();
        // entries by (lowest class, mask, first touch): groups of like masks
        std::sort(ent.begin(), ent.end(), [](const Entry& x, const Entry& y) {
            const int fx = __builtin_ctz((unsigned)x.mask), fy = __builtin_ctz((unsigned)y.mask);
            if (fx != fy) return fx < fy;
            if (x.mask != y.mask) return x.mask < y.mask;
            return x.first < y.first;
        });
        const int g_begin = (int)crow.size();
        const size_t ne = ent.size();
        n_entries += (int64_t)ne;
        for (size_t q0 = 0; q0 < ne; q0 += 16) {
            int mask = 0;
            for (size_t e = q0; e < std::min(ne, q0 + 16); ++e) mask |= ent[e].mask;
            const int nrow = 2 + __builtin_popcount((unsigned)mask);
            if (grp.size() / 64 + nrow > 0x7fffffffu) fail(EKG_EINVAL, "too many class cells");
            crow.push_back(make_int2((int)(grp.size() / 64), mask));
            const size_t base = grp.size();
            grp.resize(base + (size_t)64 * nrow);
            float* av = grp.data() + base;
            float* sq = av + 64;
            for (int lane = 0; lane < 64; ++lane) {
                const int kq = lane >> 4, b = (lane >> 2) & 3, rc = lane & 3;
                const size_t el = q0 + 4 * b + kq, ea = q0 + 4 * b + rc;
                const Entry& A = ent[ea < ne ? ea : q0];
                const bool ok = el < ne;
                const Entry& C = ent[ok ? el : q0];
                av[lane] = kq < 3 ? (float)A.c[kq] : 1.0f;
                sq[lane] = (float)(C.c[0] * C.c[0] + C.c[1] * C.c[1] + C.c[2] * C.c[2]);
                int r = 2;
                for (int i = 0; i < CLASS_WINDOW; ++i) {
                    if (!((mask >> i) & 1)) continue;
                    av[64 * r + lane] = ok ? (float)C.coef[i][rc] : 0.0f;
                    ++r;
                }
            }
        }
        const int g_end = (int)crow.size();
        // at least one piece per window, so every class has sums (0 without entries)
        const int n_pc = std::max(1, (g_end - g_begin + CLASS_PIECE - 1) / CLASS_PIECE);
        const int pc_size = std::max(1, (g_end - g_begin + n_pc - 1) / n_pc);
        int g = g_begin;
        do {
            const int ge = std::min(g_end, g + pc_size);
            pieces.push_back(int4{w, g, ge, 0});
            g = ge;
        } while (g < g_end);
        wpiece.push_back((int)pieces.size());
    }
    L.n_ccells[nhood] = n_cells;
    L.n_centries[nhood] = n_entries;
    L.n_cgroups[nhood] = (int64_t)crow.size();
    L.n_cpieces[nhood] = (int)pieces.size();
    if (grp.empty()) grp.assign(64, 0.0f);
    if (crow.empty()) crow.push_back(make_int2(0, 0));
    if (pieces.empty()) pieces.push_back(int4{0, 0, 0, 0});
    h.ctx->bind();
    L.cgrp[nhood].reserve(grp.size() * sizeof(float));
    L.crow[nhood].reserve(crow.size() * sizeof(int2));
    L.cpieces[nhood].reserve(pieces.size() * sizeof(int4));
    L.cwin[nhood].reserve(wpiece.size() * sizeof(int));
    hip_check(hipMemcpy(L.cgrp[nhood].p, grp.data(), grp.size() * sizeof(float), hipMemcpyHostToDevice),
              "upload class cell groups");
    hip_check(hipMemcpy(L.crow[nhood].p, crow.data(), crow.size() * sizeof(int2), hipMemcpyHostToDevice),
              "upload class cell group rows");
    hip_check(hipMemcpy(L.cpieces[nhood].p, pieces.data(), pieces.size() * sizeof(int4), hipMemcpyHostToDevice),
              "upload class window pieces");
    hip_check(hipMemcpy(L.cwin[nhood].p, wpiece.data(), wpiece.size() * sizeof(int), hipMemcpyHostToDevice),
              "upload class window piece offsets");
    L.cells_built[nhood] = true;
}

namespace {

// Per-request sizes of a group's packed arrays.
struct ReqSizes {
    int64_t rows, chunks, mixed, cchunks, bricks, classes;
};

ReqSizes req_sizes(const ReqSpec& r, bool mfma)
{
    const EcgLayout& L = r.h->layout;
    // the chunk kernel (matrix-core path) keeps G rows on chip: no records;
    // per-chunk sums for the chunks outside classes and per (class-window
    // piece, class of the window), partial rows for chunks outside classes
    const int64_t pieces = mfma ? L.n_cpieces[r.nhood] : 0;
    return ReqSizes{mfma ? 0 : L.V + ekgk::MFMA_PAD, (int64_t)L.n_chunks + pieces * CLASS_WINDOW, L.n_chunks,
                    pieces, L.n_bricks, L.n_classes};
}

template <class T>
void append(std::vector<char>& buf, const T* p, size_t n)
{
    const char* c = reinterpret_cast<const char*>(p);
    buf.insert(buf.end(), c, c + n * sizeof(T));
}

}  // namespace

size_t group_device_bytes(const GroupPlan& g)
{
    size_t s = 0;
    const int LP = g.LP;
    for (const ReqSpec& r : g.reqs) {
        const EcgLayout& L = r.h->layout;
        const int64_t V = L.valid ? L.V : r.h->n_myo;
        // a layout for the other path is sized as one without classes
        const bool fits = L.valid && (L.class_min == 0 || g.mfma);
        const int64_t est = std::max<int64_t>(1, V / 64 + r.h->max_layer + 4);
        const int64_t nmix = fits ? L.n_chunks : est;
        // sums rows: the chunks outside classes and, on the matrix-core path,
        // CLASS_WINDOW per class-window piece (exact once the cells are
        // built; before that bounded by a piece per window and one per
        // CLASS_PIECE groups of at most 9 entries per class voxel)
        int64_t pieces = 0;
        if (g.mfma && fits && L.n_classes > 0) {
            if (L.cells_built[g.nhood]) {
                pieces = L.n_cpieces[g.nhood];
            } else {
                int64_t cv = 0;
                for (const int4& cc : L.h_cchunks) cv += cc.z - cc.y;
                pieces = (L.n_classes + CLASS_WINDOW - 1) / CLASS_WINDOW + (9 * cv / 16 + CLASS_PIECE - 1) / CLASS_PIECE;
            }
        }
        const int64_t nall = fits ? (int64_t)L.n_chunks + pieces * CLASS_WINDOW : est;
        const int64_t ncls = fits ? L.n_classes : 0;
        const int64_t rows = g.mfma ? 0 : V + ekgk::MFMA_PAD;
        s += (size_t)r.nl * 9 * 8 + (size_t)r.L * 3 * 8 + (size_t)r.Tp * 16;
        s += (size_t)r.nl * (sizeof(ekgk::LayerConst) + r.Tp * 16);
        s += (size_t)rows * (16 * 8 + 16);
        s += (size_t)nall * (3 * LP * 8 + 8) + (size_t)ncls * 3 * LP * 8;
        s += (size_t)r.nl * ((r.Tp >> (g.mfma ? ekgk::BT_SHIFT_MFMA : ekgk::BT_SHIFT_VALU)) + 1) *
             (8 + 2 * LP * 8);
        s += (size_t)nmix * LP * r.Tp * 8;
        s += (size_t)r.L * r.T * 8;
        s += sizeof(ekgk::ReqDev) + sizeof(ekgk::HeartDev) + 9 * 8;  // record, heart, nine start tables
    }
    return s;
}

void enqueue_group(Ctx& ctx, GroupPlan& g, Workspace& ws, hipStream_t stream, const double* d_k,
                   const double* d_pos, double* d_ecg, bool time_fused, bool allow_graph)
{
    const int B = (int)g.reqs.size();
    // the launches below report hipGetLastError: drop an error that an
    // unchecked call (a destroy, a best-effort synchronise) left before them
    (void)hipGetLastError();
    if (ws.arena) {
        // a fixed workspace is shared by every call on the device: nothing of
        // an earlier group (tables, time grids) can be assumed still there
        ws.arena_used = 0;
        ws.h_tables.clear();
        ws.h_grids.clear();
    }
    const int LP = g.LP;
    const bool dev_in = d_k != nullptr;
    // hearts of the group
    std::vector<Heart*> hs;
    std::vector<ekgk::HeartDev> hdev;
    std::vector<ekgk::ReqDev> rd(B);
    // time grids shared by requests with the same start, step and length:
    // t_j = start + j*dt by repeated addition (simulator.cpp:496-500),
    // continued over the padding so the last tile stays monotone
    ws.h_tgrid.clear();
    ws.h_thl.clear();
    std::vector<std::pair<std::pair<double, double>, std::pair<int64_t, int64_t>>> grids;  // (start, dt) -> (Tp, off)
    std::vector<int64_t> st_red(B + 1, 0);
    std::vector<int64_t> st_lf(B + 1, 0), st_ch(B + 1, 0), st_xy(B + 1, 0), st_lay(B + 1, 0), st_out(B + 1, 0),
        st_cc(B + 1, 0), st_cs(B + 1, 0), st_cl(B + 1, 0);
    int64_t k_off = 0, pos_off = 0, rows = 0, c_off = 0, p_off = 0, pre_off = 0;
    if (!dev_in) {
        ws.h_k.clear();
        ws.h_pos.clear();
    }
    for (int r = 0; r < B; ++r) {
        ReqSpec& q = g.reqs[r];
        Heart& h = *q.h;
        ensure_layout(h, g.mfma);
        if (g.mfma) ensure_cells(h, g.nhood);
        int hi = -1;
        for (size_t i = 0; i < hs.size(); ++i)
            if (hs[i] == &h) hi = (int)i;
        if (hi < 0) {
            hi = (int)hs.size();
            hs.push_back(&h);
            ekgk::HeartDev d = h.layout.dev;
            d.cgrp = g.mfma ? h.layout.cgrp[g.nhood].as<float>() : nullptr;
            d.crow = g.mfma ? h.layout.crow[g.nhood].as<int2>() : nullptr;
            d.cpieces = g.mfma ? h.layout.cpieces[g.nhood].as<int4>() : nullptr;
            d.cwin = g.mfma ? h.layout.cwin[g.nhood].as<int>() : nullptr;
            hdev.push_back(d);
        }
        int64_t t_off = -1;
        for (auto& gr : grids)
            if (gr.first.first == q.start && gr.first.second == q.dt && gr.second.first >= q.Tp) t_off = gr.second.second;
        if (t_off < 0) {
            t_off = (int64_t)ws.h_tgrid.size();
            grids.push_back({{q.start, q.dt}, {q.Tp, t_off}});
            double st = q.start;
            for (int64_t j = 0; j < q.Tp; ++j) {
                ws.h_tgrid.push_back(st);
                const float hi32 = (float)st;
                ws.h_thl.push_back(hi32);
                ws.h_thl.push_back((float)(st - (double)hi32));
                st += q.dt;
            }
        }
        const ReqSizes z = req_sizes(q, g.mfma);
        ekgk::ReqDev& R = rd[r];
        R.heart = hi;
        R.L = q.L;
        R.nl = q.nl;
        R.T = (int)q.T;
        R.Tn = (int)q.Tn;
        R.Tp = (int)q.Tp;
        R.n_tiles = q.n_tiles;
        R.bt_shift = g.mfma ? ekgk::BT_SHIFT_MFMA : ekgk::BT_SHIFT_VALU;
        R.k_off = dev_in ? (int64_t)r * q.nl * 9 : k_off;
        R.pos_off = dev_in ? (int64_t)r * q.L * 3 : pos_off;
        R.ecg_off = st_out[r];
        R.t_off = t_off;
        R.lay_off = st_lay[r];
        R.xy_off = st_xy[r];
        R.g_off = rows;
        R.c_off = c_off;
        R.p_off = p_off;
        R.pre_off = pre_off;
        R.cls_off = st_cs[r];
        if (!dev_in) {
            ws.h_k.insert(ws.h_k.end(), q.k, q.k + 9 * q.nl);
            ws.h_pos.insert(ws.h_pos.end(), q.pos, q.pos + 3 * q.L);
            k_off += 9 * q.nl;
            pos_off += 3 * q.L;
        }
        rows += z.rows;
        c_off += z.chunks;
        p_off += z.mixed * LP * q.Tp;
        pre_off += (int64_t)q.nl * ((q.Tp >> R.bt_shift) + 1);
        st_lf[r + 1] = st_lf[r] + z.bricks;
        st_ch[r + 1] = st_ch[r] + z.mixed;
        st_cc[r + 1] = st_cc[r] + z.cchunks;
        st_xy[r + 1] = st_xy[r] + (int64_t)q.nl * q.Tp;
        st_lay[r + 1] = st_lay[r] + q.nl;
        st_out[r + 1] = st_out[r] + (int64_t)q.L * q.T;
        {
            const int64_t bt = (int64_t)1 << R.bt_shift;
            st_red[r + 1] = st_red[r] + (int64_t)q.L * ((q.T + bt - 1) / bt) * bt;
        }
        st_cs[r + 1] = st_cs[r] + z.classes;
        st_cl[r + 1] = st_cl[r] + (z.classes > 0 ? q.n_tiles : 0);
    }
    if (st_lf[B] > 0x7fffffffLL || st_ch[B] > 0x7fffffffLL || st_lay[B] > 0x7fffffffLL || st_cc[B] > 0x7fffffffLL ||
        st_cl[B] > 0x7fffffffLL)
        fail(EKG_EINVAL, "launch group too large");
    // device buffers (grow-only)
    ws.need(ws.lc, (size_t)st_lay[B] * sizeof(ekgk::LayerConst));
    ws.need(ws.xy, (size_t)std::max<int64_t>(1, st_xy[B]) * 16);
    ws.need(ws.G, (size_t)std::max<int64_t>(1, rows) * 16 * 8);
    ws.need(ws.aux, (size_t)std::max<int64_t>(1, rows) * sizeof(float4));
    ws.need(ws.sums, (size_t)std::max<int64_t>(1, c_off) * 3 * LP * 8);
    ws.need(ws.bnd, (size_t)std::max<int64_t>(1, c_off) * sizeof(int2));
    ws.need(ws.pre_rng, (size_t)std::max<int64_t>(1, pre_off) * sizeof(int2));
    ws.need(ws.pre_ab, (size_t)std::max<int64_t>(1, pre_off) * 2 * LP * 8);
    ws.need(ws.partial, (size_t)std::max<int64_t>(1, p_off) * 8);
    ws.need(ws.csums, (size_t)std::max<int64_t>(1, st_cs[B]) * 3 * LP * 8);
    if (!d_ecg) {
        ws.need(ws.ecg, (size_t)std::max<int64_t>(1, st_out[B]) * 8);
        d_ecg = ws.ecg.as<double>();
    }
    // class-window blocks: runs of consecutive requests of one heart (a
    // server group is ordered by heart), piece-major within a run
    std::vector<int4> runs;
    std::vector<int64_t> st_run(1, 0);
    if (g.mfma && st_cc[B] > 0) {
        for (int r = 0; r < B;) {
            int r1 = r + 1;
            while (r1 < B && g.reqs[r1].h == g.reqs[r].h) ++r1;
            const int64_t np = st_cc[r + 1] - st_cc[r];
            runs.push_back(int4{r, r1 - r, (int)np, 0});
            st_run.push_back(st_run.back() + np * (r1 - r));
            r = r1;
        }
    }
    // tables: uploaded only when they changed since the last group on this
    // workspace (a repeated group replays without host traffic)
    std::vector<char> tables;
    append(tables, hdev.data(), hdev.size());
    const size_t o_req = tables.size();
    append(tables, rd.data(), rd.size());
    const size_t o_st = tables.size();
    for (auto* st : {&st_lf, &st_ch, &st_xy, &st_lay, &st_out, &st_cc, &st_cs, &st_cl, &st_red})
        append(tables, st->data(), st->size());
    const size_t o_rst = tables.size();
    append(tables, st_run.data(), st_run.size());
    tables.resize((tables.size() + 15) & ~(size_t)15);  // int4 alignment
    const size_t o_runs = tables.size();
    append(tables, runs.data(), runs.size());
    void* before = ws.hearts.p;
    ws.need(ws.hearts, tables.size());
    if (ws.hearts.p != before) ws.h_tables.clear();
    if (tables != ws.h_tables) {
        hip_check(hipMemcpyAsync(ws.hearts.p, tables.data(), tables.size(), hipMemcpyHostToDevice, stream),
                  "upload group tables");
        ws.h_tables.swap(tables);
    }
    std::vector<char> grid_bytes;
    append(grid_bytes, ws.h_tgrid.data(), ws.h_tgrid.size());
    append(grid_bytes, ws.h_thl.data(), ws.h_thl.size());
    before = ws.tgrid.p;
    ws.need(ws.tgrid, grid_bytes.size());
    if (ws.tgrid.p != before) ws.h_grids.clear();
    if (grid_bytes != ws.h_grids) {
        hip_check(hipMemcpyAsync(ws.tgrid.p, grid_bytes.data(), grid_bytes.size(), hipMemcpyHostToDevice, stream),
                  "upload time grids");
        ws.h_grids.swap(grid_bytes);
    }
    if (!dev_in) {
        ws.need(ws.layer_k, ws.h_k.size() * 8);
        ws.need(ws.lead_pos, ws.h_pos.size() * 8);
        hip_check(hipMemcpyAsync(ws.layer_k.p, ws.h_k.data(), ws.h_k.size() * 8, hipMemcpyHostToDevice, stream),
                  "upload AP coefficients");
        hip_check(hipMemcpyAsync(ws.lead_pos.p, ws.h_pos.data(), ws.h_pos.size() * 8, hipMemcpyHostToDevice,
                                 stream),
                  "upload lead positions");
        d_k = ws.layer_k.as<double>();
        d_pos = ws.lead_pos.as<double>();
    }
    const char* tb = static_cast<const char*>(ws.hearts.p);
    const int64_t* stp = reinterpret_cast<const int64_t*>(tb + o_st);
    ekgk::GroupArgs a{};
    a.hearts = reinterpret_cast<const ekgk::HeartDev*>(tb);
    a.reqs = reinterpret_cast<const ekgk::ReqDev*>(tb + o_req);
    a.B = B;
    a.LP = LP;
    a.nhood = g.nhood;
    a.mfma = g.mfma ? 1 : 0;
    a.mfma_waves = g.waves;
    a.lf_phase = g.lf_phase;
    a.chunk_max = 0;
    for (Heart* hp : hs) a.chunk_max = std::max(a.chunk_max, hp->layout.chunk_target);
    a.lf_start = stp;
    a.ch_start = stp + (B + 1);
    a.xy_start = stp + 2 * (B + 1);
    a.lay_start = stp + 3 * (B + 1);
    a.out_start = stp + 4 * (B + 1);
    a.cs_start = stp + 6 * (B + 1);
    a.cl_start = stp + 7 * (B + 1);
    a.red_start = stp + 8 * (B + 1);
    a.n_cc = st_cc[B];
    a.n_cs = st_cs[B];
    a.cs_max = 0;
    for (int r = 0; r < B; ++r) a.cs_max = std::max<int>(a.cs_max, (int)(st_cs[r + 1] - st_cs[r]));
    a.n_cl = st_cl[B];
    a.n_lf = st_lf[B];
    a.n_ch = st_ch[B];
    // every request with the same chunk count: the chunk kernel runs chunk-major
    a.ch_uniform = B > 0 ? (int)(st_ch[1] - st_ch[0]) : 0;
    for (int r = 1; r < B && a.ch_uniform; ++r)
        if (st_ch[r + 1] - st_ch[r] != a.ch_uniform) a.ch_uniform = 0;
    a.one_heart = hs.size() == 1 ? 1 : 0;
    a.cc_runs = reinterpret_cast<const int4*>(tb + o_runs);
    a.cc_rstart = reinterpret_cast<const int64_t*>(tb + o_rst);
    a.n_runs = (int)runs.size();
    // class-kernel blocks: XCD-aware order when every request has the same count
    a.cl_per_req = (B > 0 && B % 8 == 0) ? (int)(st_cl[1] - st_cl[0]) : 0;
    for (int r = 1; r < B && a.cl_per_req; ++r)
        if (st_cl[r + 1] - st_cl[r] != a.cl_per_req) a.cl_per_req = 0;
    a.n_xy = st_xy[B];
    a.n_lay = st_lay[B];
    a.n_out = st_out[B];
    a.n_red = st_red[B];
    a.layer_k = d_k;
    a.lead_pos = d_pos;
    a.tgrid = ws.tgrid.as<double>();
    a.thl = reinterpret_cast<const float2*>(static_cast<const char*>(ws.tgrid.p) + ws.h_tgrid.size() * 8);
    a.lc = ws.lc.as<ekgk::LayerConst>();
    a.xy = ws.xy.as<double2>();
    a.rec = ws.G.as<double>();
    a.aux = ws.aux.as<float4>();
    a.sums = ws.sums.as<double>();
    a.bnd = ws.bnd.as<int2>();
    a.pre_rng = ws.pre_rng.as<int2>();
    a.pre_ab = ws.pre_ab.as<double>();
    a.partial = ws.partial.as<double>();
    a.csums = ws.csums.as<double>();
    a.ecg = d_ecg;
    auto launch_all = [&](bool timed) {
        // The timed kernel is the dominant one of each path: the class-window
        // kernel, or without classes the chunk kernel (matrix-core path), or
        // the fused transition-tile kernel (vector ALUs).
        hipEvent_t e0 = nullptr, e1 = nullptr;
        auto tic = [&]() {
            if (!timed) return;
            e0 = ctx.get_event();
            e1 = ctx.get_event();
            hip_check(hipEventRecord(e0, stream), "hipEventRecord");
        };
        auto toc = [&]() {
            if (!timed) return;
            hip_check(hipEventRecord(e1, stream), "hipEventRecord");
            ctx.pending.emplace_back(e0, e1);
        };
        if (g.mfma && a.n_cc > 0 && ws.side) {
            // Two branches: the class-window kernel and the class sums on
            // `stream`, beside the AP tables, chunk, prefix and reduction
            // kernels on the side stream (they share no buffer: the chunk
            // and class-window kernels write disjoint rows of the sums); the
            // class ECG kernel adds to the reduction's output after both.
            hip_check(hipEventRecord(ws.ev_fork, stream), "hipEventRecord");
            hip_check(hipStreamWaitEvent(ws.side, ws.ev_fork, 0), "hipStreamWaitEvent");
            hip_check(ekgk::launch_ap_tables(a, ws.side), "ap_tables launch");
            hip_check(ekgk::launch_ecg_chunk(a, ws.side), "ecg_chunk launch");
            hip_check(ekgk::launch_layer_prefix(a, ws.side), "layer_prefix launch");
            hip_check(ekgk::launch_ecg_reduce(a, ws.side), "ecg_reduce launch");
            hip_check(hipEventRecord(ws.ev_join, ws.side), "hipEventRecord");
            tic();
            hip_check(ekgk::launch_class_windows(a, stream), "class_window launch");
            toc();
            hip_check(ekgk::launch_class_sums(a, stream), "class_sums launch");
            hip_check(hipStreamWaitEvent(stream, ws.ev_join, 0), "hipStreamWaitEvent");
            hip_check(ekgk::launch_ecg_class(a, stream), "ecg_class launch");
            return;
        }
        hip_check(ekgk::launch_ap_tables(a, stream), "ap_tables launch");
        if (g.mfma) {
            // the class-window kernel is timed when the group has classes
            if (a.n_cc == 0) tic();
            hip_check(ekgk::launch_ecg_chunk(a, stream), "ecg_chunk launch");
            if (a.n_cc == 0) toc();
            if (a.n_cc > 0) tic();
            hip_check(ekgk::launch_class_windows(a, stream), "class_window launch");
            if (a.n_cc > 0) toc();
            hip_check(ekgk::launch_class_sums(a, stream), "class_sums launch");
            hip_check(ekgk::launch_layer_prefix(a, stream), "layer_prefix launch");
        } else {
            hip_check(ekgk::launch_lead_field(a, stream), "lead_field launch");
            hip_check(ekgk::launch_chunk_sums(a, stream), "chunk_sums launch");
            hip_check(ekgk::launch_layer_prefix(a, stream), "layer_prefix launch");
            tic();
            hip_check(ekgk::launch_ecg_fused(a, stream), "ecg_fused launch");
            toc();
        }
        hip_check(ekgk::launch_ecg_reduce(a, stream), "ecg_reduce launch");
        hip_check(ekgk::launch_ecg_class(a, stream), "ecg_class launch");
    };
    if (!ctx.graphs || time_fused || !allow_graph) {
        launch_all(time_fused);
        return;
    }
    // Graph mode: the launches of a group are captured once and replayed
    // while the kernel arguments (buffers, counts, shapes) stay the same; the
    // per-request values live in device tables and buffers, so new inputs
    // in the same places need no recapture.  The graph holds kernels only.
    std::vector<char> key(reinterpret_cast<const char*>(&a), reinterpret_cast<const char*>(&a) + sizeof a);
    if (!ws.graph || key != ws.graph_key) {
        if (ws.graph) {
            (void)hipGraphExecDestroy(ws.graph);
            ws.graph = nullptr;
        }
        ws.graph_key.clear();
        hip_check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        hipGraph_t graph = nullptr;
        try {
            launch_all(false);
        } catch (...) {
            (void)hipStreamEndCapture(stream, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        hip_check(hipStreamEndCapture(stream, &graph), "hipStreamEndCapture");
        const hipError_t e = hipGraphInstantiate(&ws.graph, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        hip_check(e, "hipGraphInstantiate");
        ws.graph_key.swap(key);
        ++ctx.graph_captures;
    }
    hip_check(hipGraphLaunch(ws.graph, stream), "hipGraphLaunch");
    ++ctx.graph_launches;
}

void simulate(Heart& h, int batch, const ekg_sim_desc* descs, double* ecg)
{
    Ctx& ctx = *h.ctx;
    std::lock_guard<std::mutex> g1(ctx.mu);
    std::lock_guard<std::mutex> g2(h.mu);
    if (!ecg) fail(EKG_EINVAL, "ecg output buffer is required");
    if (batch < 1 || !descs) fail(EKG_EINVAL, "batch must hold at least one descriptor");
    std::vector<ReqSpec> specs;
    for (int b = 0; b < batch; ++b) specs.push_back(make_spec(h, descs[b], true));
    int64_t n_out = 0;
    for (const ReqSpec& r : specs) n_out += (int64_t)r.L * r.T;
    if (n_out == 0) return;
    ctx.bind();
    GroupPlan g = plan_group(ctx, std::move(specs));
    enqueue_group(ctx, g, ctx.ws, ctx.stream, nullptr, nullptr, nullptr, ctx.profiling);
    hip_check(hipMemcpyAsync(ecg, ctx.ws.ecg.p, (size_t)n_out * 8, hipMemcpyDeviceToHost, ctx.stream),
              "download ECG");
    hip_check(hipStreamSynchronize(ctx.stream), "hipStreamSynchronize");
    if (ctx.profiling) ctx.collect_timings();
}

void simulate_device(Heart& h, const ekg_batch_dev& d)
{
    Ctx& ctx = *h.ctx;
    std::lock_guard<std::mutex> g1(ctx.mu);
    std::lock_guard<std::mutex> g2(h.mu);
    if (!d.layer_k_dev || !d.lead_pos_dev || !d.ecg_dev) fail(EKG_EINVAL, "device pointers are required");
    if (d.batch < 1 || d.batch > (1 << 20)) fail(EKG_EINVAL, "batch must hold 1..2^20 simulations");
    ekg_sim_desc s{};
    s.n_layers = d.n_layers;
    s.n_leads = d.n_leads;
    s.nhood = d.nhood;
    s.start = d.start;
    s.length = d.length;
    s.dt = d.dt;
    const ReqSpec one = make_spec(h, s, false);
    if (one.T == 0) return;
    ctx.bind();
    GroupPlan g = plan_group(ctx, std::vector<ReqSpec>(d.batch, one));
    enqueue_group(ctx, g, ctx.ws, ctx.stream, d.layer_k_dev, d.lead_pos_dev, d.ecg_dev, ctx.profiling);
}

void simulate_device_fixed(Heart& h, const ekg_batch_dev& d)
{
    Ctx& ctx = *h.ctx;
    std::lock_guard<std::mutex> g1(ctx.mu);
    std::lock_guard<std::mutex> g2(h.mu);
    if (!d.layer_k_dev || !d.lead_pos_dev || !d.ecg_dev) fail(EKG_EINVAL, "device pointers are required");
    if (d.batch < 1 || d.batch > (1 << 20)) fail(EKG_EINVAL, "batch must hold 1..2^20 simulations");
    if (!ctx.ws_fixed.arena) fail(EKG_ESTATE, "the device's static workspace is not resolved");
    ekg_sim_desc s{};
    s.n_layers = d.n_layers;
    s.n_leads = d.n_leads;
    s.nhood = d.nhood;
    s.start = d.start;
    s.length = d.length;
    s.dt = d.dt;
    const ReqSpec one = make_spec(h, s, false);
    if (one.T == 0) return;
    GroupPlan g = plan_group(ctx, std::vector<ReqSpec>(d.batch, one));
    // the per-heart layout is device memory built on first use: it must exist
    if (!layout_ready(h, g.mfma, g.nhood))
        fail(EKG_ESTATE, "the heart's layout for this neighbourhood and tuning is not built; call "
                         "ekg_heart_layout_info (or any simulation of the heart) first");
    ctx.bind();
    enqueue_group(ctx, g, ctx.ws_fixed, ctx.stream, d.layer_k_dev, d.lead_pos_dev, d.ecg_dev, false, false);
}

ekg_tune autotune(Heart& h, const ekg_sim_desc& sample, int batch)
{
    Ctx& ctx = *h.ctx;
    std::lock_guard<std::mutex> g1(ctx.mu);
    std::lock_guard<std::mutex> g2(h.mu);
    if (batch < 1 || batch > 4096) fail(EKG_EINVAL, "autotune batch must be 1..4096");
    const ReqSpec one = make_spec(h, sample, true);
    if (one.T == 0) fail(EKG_EINVAL, "the sample simulation has no time steps");
    if (!h.excited) fail(EKG_ESTATE, "excitation sequence not computed or loaded");
    ctx.bind();
    const int saved_kernel = ctx.kernel;
    const int def_chunk = DEFAULT_CHUNK;
    ekg_tune best{def_chunk, DEFAULT_WAVES, 2, EKG_KERNEL_MFMA, h.tune_class, 0.0};
    Workspace ws;
    hipStream_t st = ctx.stream;
    hipEvent_t e0 = ctx.get_event(), e1 = ctx.get_event();
    // time one configuration: one warm-up group, then the best of two
    auto trial = [&](const ekg_tune& c) {
        h.tune_chunk = c.chunk_voxels;
        h.tune_waves = c.mfma_waves;
        h.tune_phase = c.lead_phase;
        h.tune_kernel = c.kernel;
        h.tune_class = c.class_min;
        ctx.kernel = EKG_KERNEL_AUTO;
        GroupPlan g = plan_group(ctx, std::vector<ReqSpec>(batch, one));
        double best_ms = 1e30;
        for (int it = 0; it < 3; ++it) {
            hip_check(hipEventRecord(e0, st), "hipEventRecord");
            enqueue_group(ctx, g, ws, st, nullptr, nullptr, nullptr, false);
            hip_check(hipEventRecord(e1, st), "hipEventRecord");
            hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
            float ms = 0.f;
            hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
            if (it > 0) best_ms = std::min(best_ms, (double)ms);
        }
        return best_ms;
    };
    try {
        best.ms_per_batch = trial(best);
        auto sweep = [&](int ekg_tune::*field, std::vector<int> values) {
            for (int v : values) {
                if (v == best.*field) continue;
                ekg_tune c = best;
                c.*field = v;
                const double ms = trial(c);
                if (ms < best.ms_per_batch) {
                    best = c;
                    best.ms_per_batch = ms;
                }
            }
        };
        sweep(&ekg_tune::kernel, {EKG_KERNEL_MFMA, EKG_KERNEL_VALU});
        if (best.kernel == EKG_KERNEL_MFMA) {
            // chunk size and block size interact (LDS rows per block against
            // blocks per CU): every pair
            const ekg_tune base = best;
            for (int chunk : {64, 128, 256, 512})
                for (int waves : {2, 4, 8}) {
                    if (chunk == base.chunk_voxels && waves == base.mfma_waves) continue;
                    ekg_tune c = base;
                    c.chunk_voxels = chunk;
                    c.mfma_waves = waves;
                    const double ms = trial(c);
                    if (ms < best.ms_per_batch) {
                        best = c;
                        best.ms_per_batch = ms;
                    }
                }
            // every voxel in a class (no chunks outside classes) when the
            // default threshold already leaves at most 1% of the voxels out:
            // the few extra classes are cheaper than the chunk path for them
            if (best.class_min == 0 && h.layout.valid && h.layout.n_classes > 0) {
                int64_t cv = 0;
                for (const int4& cc : h.layout.h_cchunks) cv += cc.z - cc.y;
                if (h.layout.V - cv <= h.layout.V / 100) sweep(&ekg_tune::class_min, {1});
            }
        } else {
            std::vector<int> chunks{64, 128, 256, 512};
            chunks.erase(std::remove(chunks.begin(), chunks.end(), def_chunk), chunks.end());
            sweep(&ekg_tune::chunk_voxels, chunks);
            sweep(&ekg_tune::lead_phase, {2, 4});
        }
    } catch (...) {
        ctx.kernel = saved_kernel;
        h.tune_chunk = h.tune_waves = h.tune_phase = h.tune_kernel = 0;
        ctx.free_events.push_back(e0);
        ctx.free_events.push_back(e1);
        throw;
    }
    ctx.kernel = saved_kernel;
    ctx.free_events.push_back(e0);
    ctx.free_events.push_back(e1);
    h.tune_chunk = best.chunk_voxels;
    h.tune_waves = best.mfma_waves;
    h.tune_phase = best.lead_phase;
    h.tune_kernel = best.kernel;
    h.tune_class = best.class_min;
    ensure_layout(h, best.kernel != EKG_KERNEL_VALU);
    hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
    return best;
}

namespace {

int crit_count(const ekg_criteria_desc& c, int n_leads)
{
    if (c.mode < EKG_CMP_RMS || c.mode > EKG_CMP_VECTOR_CORRELATION) fail(EKG_EINVAL, "unknown comparison mode");
    if (c.n_targets < 1 || c.target_len < 1 || !c.targets || !c.target_offsets)
        fail(EKG_EINVAL, "criteria need at least one non-empty target and its offset");
    return std::min(n_leads, c.n_targets);
}

// Targets and offsets into the context's buffers on its stream, and their
// statistics over the comparison length for ECGs of n_steps samples -- only
// when these differ from the last call (an optimiser keeps its targets).
void prepare_targets(Ctx& ctx, const ekg_criteria_desc& c, int64_t n_steps)
{
    const size_t nt = (size_t)c.n_targets * (size_t)c.target_len;
    const int64_t len = std::min<int64_t>(n_steps, c.target_len);
    std::vector<double> h(c.targets, c.targets + nt);
    h.insert(h.end(), c.target_offsets, c.target_offsets + c.n_targets);
    h.push_back((double)c.n_targets);
    h.push_back((double)len);
    void* b0 = ctx.crit_tg.p;
    void* b1 = ctx.crit_off.p;
    void* b2 = ctx.crit_stats.p;
    ctx.crit_tg.reserve(nt * 8);
    ctx.crit_off.reserve((size_t)c.n_targets * 8);
    ctx.crit_stats.reserve((size_t)c.n_targets * 4 * 8);
    if (ctx.crit_tg.p != b0 || ctx.crit_off.p != b1 || ctx.crit_stats.p != b2) ctx.crit_h.clear();
    if (h == ctx.crit_h) return;
    hip_check(hipMemcpyAsync(ctx.crit_tg.p, c.targets, nt * 8, hipMemcpyHostToDevice, ctx.stream), "upload targets");
    hip_check(hipMemcpyAsync(ctx.crit_off.p, c.target_offsets, (size_t)c.n_targets * 8, hipMemcpyHostToDevice,
                             ctx.stream),
              "upload target offsets");
    hip_check(ekgk::launch_target_stats(ctx.crit_tg.as<double>(), c.n_targets, c.target_len, len,
                                        ctx.crit_stats.as<double>(), ctx.stream),
              "target statistics launch");
    ctx.crit_h.swap(h);
}

}  // namespace

void criteria_device(Ctx& ctx, const ekg_criteria_desc& c, const double* d_ecg, int batch, int n_leads,
                     int64_t n_steps, double* d_out)
{
    std::lock_guard<std::mutex> g(ctx.mu);
    if (batch < 0 || n_leads < 1 || n_steps < 1) fail(EKG_EINVAL, "bad batch shape");
    if (!d_ecg || !d_out) fail(EKG_EINVAL, "NULL device buffer");
    const int n_crit = crit_count(c, n_leads);
    if (batch == 0) return;
    ctx.bind();
    prepare_targets(ctx, c, n_steps);
    hip_check(ekgk::launch_criteria(c.mode, c.peaks ? 1 : 0, c.raw ? 1 : 0, d_ecg, batch, n_leads, n_steps,
                                    ctx.crit_tg.as<double>(), ctx.crit_off.as<double>(), ctx.crit_stats.as<double>(),
                                    n_crit, c.target_len, d_out, ctx.stream),
              "criteria launch");
}

void simulate_criteria(Heart& h, int batch, const ekg_sim_desc* descs, const ekg_criteria_desc& c,
                       double* criteria)
{
    Ctx& ctx = *h.ctx;
    std::lock_guard<std::mutex> g1(ctx.mu);
    std::lock_guard<std::mutex> g2(h.mu);
    if (!criteria) fail(EKG_EINVAL, "criteria output buffer is required");
    if (batch < 1 || !descs) fail(EKG_EINVAL, "batch must hold at least one descriptor");
    std::vector<ReqSpec> specs;
    for (int b = 0; b < batch; ++b) specs.push_back(make_spec(h, descs[b], true));
    const int L = specs[0].L;
    const int64_t T = specs[0].T;
    for (const ReqSpec& r : specs)
        if (r.L != L || r.T != T) fail(EKG_EINVAL, "every request of a criteria batch needs the same leads and steps");
    if (T < 1) fail(EKG_EINVAL, "no time steps");
    const int n_crit = crit_count(c, L);
    const size_t per = (size_t)n_crit * (c.peaks ? 2 : 1);
    ctx.bind();
    GroupPlan g = plan_group(ctx, std::move(specs));
    enqueue_group(ctx, g, ctx.ws, ctx.stream, nullptr, nullptr, nullptr, ctx.profiling);
    prepare_targets(ctx, c, T);
    ctx.crit_out.reserve((size_t)batch * per * 8);
    hip_check(ekgk::launch_criteria(c.mode, c.peaks ? 1 : 0, c.raw ? 1 : 0, ctx.ws.ecg.as<double>(), batch, L, T,
                                    ctx.crit_tg.as<double>(), ctx.crit_off.as<double>(), ctx.crit_stats.as<double>(),
                                    n_crit, c.target_len, ctx.crit_out.as<double>(), ctx.stream),
              "criteria launch");
    hip_check(hipMemcpyAsync(criteria, ctx.crit_out.p, (size_t)batch * per * 8, hipMemcpyDeviceToHost, ctx.stream),
              "download criteria");
    hip_check(hipStreamSynchronize(ctx.stream), "hipStreamSynchronize");
    if (ctx.profiling) ctx.collect_timings();
}

void excite(Heart& h, bool is_int, const double* transfer, int rows, int cols, const uint8_t* delay)
{
    Ctx& ctx = *h.ctx;
    std::lock_guard<std::mutex> g1(ctx.mu);
    std::lock_guard<std::mutex> g2(h.mu);
    if (h.h_seed.empty()) fail(EKG_ENOSEED, "Could not find starting point for excitation sequence");
    int64_t nseeds = 0;
    for (int64_t i = 0; i < h.nvox; ++i)
        if (h.h_seed[i]) {
            if (!h.h_layer[i]) fail(EKG_EINVAL, "excitation starting point on an empty voxel");
            ++nseeds;
        }
    if (nseeds == 0) fail(EKG_ENOSEED, "Could not find starting point for excitation sequence");
    std::vector<double> wtab;
    if (!is_int) {
        if (!transfer || rows < 1 || cols < 1) fail(EKG_EINVAL, "transfer matrix required");
        if (h.max_layer >= cols || h.max_layer >= rows)
            fail(EKG_ETRANSFER, "transfer (conduction) matrix does not define layer " +
                                    std::to_string(h.max_layer));
        for (int a = 1; a <= h.max_layer; ++a)
            for (int b = 1; b <= h.max_layer; ++b) {
                const double t = transfer[a * cols + b];
                if (!(t >= 0) || !std::isfinite(t))
                    fail(EKG_ETRANSFER, "conduction matrix entries between tissue layers must be "
                                        "finite and >= 0");
            }
        // lag = transfer[lu][lv] * sqrt(|d|^2), FP64 (simulator.cpp:239-240)
        wtab.assign((size_t)rows * cols * 4, 0.0);
        for (int a = 0; a < rows; ++a)
            for (int b = 0; b < cols; ++b)
                for (int k = 1; k <= 3; ++k) {
                    double lag = transfer[a * cols + b];
                    lag *= std::sqrt((double)k);
                    wtab[((size_t)a * cols + b) * 4 + k] = lag;
                }
    } else if (!delay) {
        fail(EKG_EINVAL, "delay grid required");
    }
    ctx.bind();
    const int nb0 = (h.n0 + 7) / 8, nb1 = (h.n1 + 7) / 8, nb2 = (h.n2 + 7) / 8;
    const int nbricks = nb0 * nb1 * nb2;
    // Sweeps are enqueued in rounds of growing length (4, 8, 16, ... up to
    // 64) with one flag per sweep; the host reads the flags once per round.
    // A sweep after convergence finds no active brick and changes nothing.
    constexpr int MAX_ROUND = 64;
    DevBuf d_dist, d_wtab, d_delay, d_flags, d_any;
    d_dist.reserve((size_t)h.nvox * (is_int ? 4 : 8));
    d_flags.reserve((size_t)nbricks * 2);
    d_any.reserve(MAX_ROUND * sizeof(int));
    if (!is_int) {
        d_wtab.reserve(wtab.size() * 8);
        hip_check(hipMemcpy(d_wtab.p, wtab.data(), wtab.size() * 8, hipMemcpyHostToDevice),
                  "upload lag table");
    } else {
        d_delay.reserve(h.nvox);
        hip_check(hipMemcpy(d_delay.p, delay, h.nvox, hipMemcpyHostToDevice), "upload delays");
    }
    uint8_t* fa = d_flags.as<uint8_t>();
    uint8_t* fb = fa + nbricks;
    hipStream_t s = ctx.stream;
    hip_check(hipMemsetAsync(fa, 0, (size_t)nbricks * 2, s), "hipMemsetAsync");
    hip_check(ekgk::launch_ca_init(is_int, h.d_layer.as<uint8_t>(), h.d_seed.as<uint8_t>(), d_dist.p,
                                   fa, h.n0, h.n1, h.n2, nb1, nb2, s),
              "ca_init launch");
    int64_t sweeps = 0;
    const int64_t max_sweeps = 64 + 8 * (int64_t)h.nvox;
    int flags[MAX_ROUND];
    for (int round = 4;; round = std::min(2 * round, MAX_ROUND)) {
        hip_check(hipMemsetAsync(d_any.p, 0, (size_t)round * sizeof(int), s), "hipMemsetAsync");
        for (int q = 0; q < round; ++q) {
            hip_check(hipMemsetAsync(fb, 0, nbricks, s), "hipMemsetAsync");
            hip_check(ekgk::launch_ca_sweep(is_int, h.d_layer.as<uint8_t>(), d_delay.as<uint8_t>(),
                                            d_wtab.as<double>(), cols, d_dist.p, h.n0, h.n1, h.n2, nb0,
                                            nb1, nb2, fa, fb, d_any.as<int>() + q, s),
                      "ca_sweep launch");
            std::swap(fa, fb);
        }
        hip_check(hipMemcpyAsync(flags, d_any.p, (size_t)round * sizeof(int), hipMemcpyDeviceToHost, s),
                  "read flags");
        hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
        sweeps += round;
        if (!flags[round - 1]) break;  // the round's last sweep marked no brick: converged
        if (sweeps > max_sweeps) fail(EKG_EDEVICE, "excitation relaxation did not converge");
    }
    h.d_act.reserve((size_t)h.nvox * 8);
    hip_check(ekgk::launch_ca_finalize(is_int, d_dist.p, h.d_act.as<double>(), h.nvox, s),
              "ca_finalize launch");
    h.h_act.resize(h.nvox);
    hip_check(hipMemcpyAsync(h.h_act.data(), h.d_act.p, (size_t)h.nvox * 8, hipMemcpyDeviceToHost, s),
              "download activation");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    h.excited = true;
    h.layout.valid = false;
}

}  // namespace ekg
