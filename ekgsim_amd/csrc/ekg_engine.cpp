// ekg_engine.cpp -- host side of the plain and batched simulation paths:
// heart upload, excitation-sequence driver, ECG layout, batch planning and
// the kernel launch sequence.  All validation happens here, before any
// launch, so that every kernel runs on extents it was sized for.
#include "ekg_engine.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

namespace ekg {

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

void DevBuf::release()
{
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}

void Ctx::bind() const { hip_check(hipSetDevice(device), "hipSetDevice"); }

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

const void* ctx_nb_table(const Ctx& ctx, int nhood)
{
    return static_cast<const int8_t*>(ctx.nb_table.p) + nhood * 26 * 4;
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

// Shape checks and derived sizes shared by the host and device entry points.
BatchPlan plan_shape(const Heart& h, BatchPlan p)
{
    if (p.B < 1 || p.B > 65535) fail(EKG_EINVAL, "batch must hold 1..65535 simulations");
    if (p.L < 1 || p.L > EKG_MAX_LEADS)
        fail(EKG_EINVAL, "n_leads must be in 1.." + std::to_string(EKG_MAX_LEADS));
    if (p.nhood < EKG_NHOOD_2D4 || p.nhood > EKG_NHOOD_3D8)
        fail(EKG_EINVAL, "unknown neighbourhood (only know of these: 2D4, 3D4, 2D8, 3D8 (cube))");
    if (p.n_layers < h.max_layer || p.n_layers < 1)
        fail(EKG_EINVAL, "the heart has " + std::to_string(h.max_layer) + " layers but only " +
                             std::to_string(p.n_layers) + " layer APs were given");
    if (!(p.dt > 0) || !std::isfinite(p.dt) || !std::isfinite(p.start) || !std::isfinite(p.length))
        fail(EKG_EINVAL, "time step must be positive and times finite");
    p.T = num_steps(p.length, p.dt);
    if (p.T < 0 || p.T > (1 << 24)) fail(EKG_EINVAL, "too many time steps");
    p.LP = p.L <= 4 ? 4 : p.L <= 8 ? 8 : p.L <= 12 ? 12 : 16;
    p.Tp = ((p.T + 63) / 64) * 64;
    p.n_tiles = (int)(p.Tp / 64);
    p.n_groups = (p.n_tiles + 3) / 4;
    p.Tn = neighbour_steps(p.start, p.length, p.dt, p.T + 1);
    if (p.Tn > p.T)
        fail(EKG_EINVAL, "time step makes the neighbour loop of the reference run past the "
                         "sample buffer (simulator.cpp:519); choose another time step");
    return p;
}

// Matrix-core contraction unless the caller pinned the VALU kernel.
bool use_mfma(const Ctx& ctx, int n_leads)
{
    (void)n_leads;
    return ctx.kernel != EKG_KERNEL_VALU;
}

BatchPlan plan_batch(const Heart& h, int batch, const ekg_sim_desc* descs)
{
    if (batch < 1 || !descs) fail(EKG_EINVAL, "batch must hold at least one descriptor");
    BatchPlan p;
    const ekg_sim_desc& d0 = descs[0];
    p.B = batch;
    p.L = d0.n_leads;
    p.n_layers = d0.n_layers;
    p.nhood = d0.nhood;
    p.start = d0.start;
    p.length = d0.length;
    p.dt = d0.dt;
    for (int b = 0; b < batch; ++b) {
        const ekg_sim_desc& d = descs[b];
        if (!d.layer_k || !d.lead_pos) fail(EKG_EINVAL, "layer_k and lead_pos are required");
        if (d.n_leads != p.L || d.nhood != p.nhood || d.start != p.start || d.length != p.length ||
            d.dt != p.dt || d.n_layers != p.n_layers)
            fail(EKG_EINVAL, "all descriptors of a batch must share leads, neighbourhood, "
                             "layers, start, length and time step");
        if (d.n_layers < 1 || d.n_leads < 1 || d.n_leads > EKG_MAX_LEADS)
            fail(EKG_EINVAL, "invalid layer or lead count");
        for (int i = 0; i < 9 * d.n_layers; ++i)
            if (!std::isfinite(d.layer_k[i])) fail(EKG_EINVAL, "non-finite AP coefficient");
        for (int i = 0; i < 3 * d.n_leads; ++i)
            if (!std::isfinite(d.lead_pos[i])) fail(EKG_EINVAL, "non-finite lead position");
    }
    return plan_shape(h, p);
}

// Largest activation-time span inside one chunk (ms).  The saturated AP path
// scales exponentials by e^{(k4+k5)(a - t_ref)} per voxel; a bounded span
// keeps that factor far from FP32 overflow for any realistic k4, k5.
constexpr double CHUNK_SPAN_MS = 16.0;

void ensure_layout(Heart& h)
{
    if (!h.excited) fail(EKG_ESTATE, "excitation sequence not computed or loaded");
    EcgLayout& L = h.layout;
    if (L.valid) return;
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
    // Chunk size: enough waves for one simulation to fill the GPU
    // (~400 chunks x time tiles), bounded so that per-wave setup stays small;
    // chunks are layer-pure and span at most CHUNK_SPAN_MS of activation time.
    const int64_t target = std::min<int64_t>(4096, std::max<int64_t>(256, V / 400));
    int64_t v = 0;
    while (v < V) {
        const int lay = h.h_layer[idx[v]];
        int64_t e = v;
        while (e < V && h.h_layer[idx[e]] == lay) ++e;
        const int64_t n = e - v;
        const int64_t k = (n + target - 1) / target;
        for (int64_t c = 0; c < k; ++c) {
            int64_t b0 = v + n * c / k;
            const int64_t b1 = v + n * (c + 1) / k;
            while (b0 < b1) {
                const float tref = (float)h.h_act[idx[b0]];
                int64_t b2 = b0 + 1;
                while (b2 < b1 && h.h_act[idx[b2]] - (double)tref <= CHUNK_SPAN_MS) ++b2;
                int tbits;
                std::memcpy(&tbits, &tref, 4);
                L.h_chunks.push_back(int4{lay, (int)b0, (int)b2, tbits});
                b0 = b2;
            }
        }
        v = e;
    }
    L.n_chunks = (int)L.h_chunks.size();
    // per sorted voxel: chunk index and activation time (FP64 and hi/lo FP32)
    std::vector<int> vchunk(V ? V : 1);
    for (int c = 0; c < L.n_chunks; ++c)
        for (int v2 = L.h_chunks[c].y; v2 < L.h_chunks[c].z; ++v2) vchunk[v2] = c;
    std::vector<double> a64(V ? V : 1);
    std::vector<float2> af(V ? V : 1);
    for (int64_t i = 0; i < V; ++i) {
        const double a = h.h_act[idx[i]];
        a64[i] = a;
        const float hi = (float)a;
        af[i] = float2{hi, (float)(a - (double)hi)};
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
    if (!L.h_chunks.empty())
        hip_check(hipMemcpy(L.chunks.p, L.h_chunks.data(), L.h_chunks.size() * sizeof(int4),
                            hipMemcpyHostToDevice),
                  "upload chunks");
    L.valid = true;
}

size_t batch_device_bytes(const Heart& h, const BatchPlan& p)
{
    const int64_t V = h.layout.valid ? h.layout.V : h.n_myo;
    const int64_t nch = h.layout.valid ? h.layout.n_chunks : std::max<int64_t>(1, V / 256 + 3);
    size_t s = 0;
    s += (size_t)p.B * p.n_layers * 9 * 8;
    s += (size_t)p.B * p.L * 3 * 8;
    s += (size_t)p.Tp * (8 + 8);
    s += (size_t)p.B * p.n_layers * p.Tp * 8;
    s += (size_t)p.B * p.n_layers * sizeof(ekgk::LayerConst);
    s += p.mfma ? (size_t)p.B * (V + ekgk::MFMA_PAD) * (16 * 8 + 16)
                : (size_t)p.B * (V + 1) * (p.LP + 2) * 8;
    s += (size_t)p.B * nch * 3 * p.LP * 8;
    s += (size_t)p.B * nch * p.LP * p.Tp * 8;
    s += (size_t)p.B * p.L * p.T * 8;
    return s;
}

void upload_batch(const BatchPlan& p, const ekg_sim_desc* descs, Workspace& ws, hipStream_t stream)
{
    const int B = p.B;
    ws.h_k.resize((size_t)B * p.n_layers * 9);
    ws.h_pos.resize((size_t)B * p.L * 3);
    for (int b = 0; b < B; ++b) {
        std::memcpy(&ws.h_k[(size_t)b * p.n_layers * 9], descs[b].layer_k,
                    sizeof(double) * 9 * p.n_layers);
        std::memcpy(&ws.h_pos[(size_t)b * p.L * 3], descs[b].lead_pos, sizeof(double) * 3 * p.L);
    }
    ws.layer_k.reserve(ws.h_k.size() * 8);
    ws.lead_pos.reserve(ws.h_pos.size() * 8);
    ws.ecg.reserve((size_t)B * p.L * std::max<int64_t>(p.T, 1) * 8);
    hip_check(hipMemcpyAsync(ws.layer_k.p, ws.h_k.data(), ws.h_k.size() * 8, hipMemcpyHostToDevice,
                             stream),
              "upload AP coefficients");
    hip_check(hipMemcpyAsync(ws.lead_pos.p, ws.h_pos.data(), ws.h_pos.size() * 8,
                             hipMemcpyHostToDevice, stream),
              "upload lead positions");
}

void enqueue_batch(Ctx& ctx, Heart& h, const BatchPlan& p, const double* d_layer_k,
                   const double* d_lead_pos, double* d_ecg, Workspace& ws, hipStream_t stream,
                   bool time_fused)
{
    const EcgLayout& Lo = h.layout;
    const int B = p.B;
    const int64_t V = Lo.V;
    const int nch = Lo.n_chunks;
    // time grid t_j = start + j*dt by repeated addition (simulator.cpp:496-500),
    // continued over the padding so the last tile stays monotone
    if ((int64_t)ws.h_tgrid.size() != p.Tp || ws.h_tgrid.empty() || ws.h_tgrid[0] != p.start ||
        ws.grid_dt != p.dt) {
        ws.h_tgrid.resize(p.Tp);
        ws.h_thl.resize(2 * p.Tp);
        double st = p.start;
        for (int64_t j = 0; j < p.Tp; ++j) {
            ws.h_tgrid[j] = st;
            const float hi = (float)st;
            ws.h_thl[2 * j] = hi;
            ws.h_thl[2 * j + 1] = (float)(st - (double)hi);
            st += p.dt;
        }
        ws.grid_dt = p.dt;
        ws.tgrid.reserve(p.Tp * 8);
        ws.thl.reserve(p.Tp * 8);
        hip_check(hipMemcpyAsync(ws.tgrid.p, ws.h_tgrid.data(), p.Tp * 8, hipMemcpyHostToDevice, stream),
                  "upload time grid");
        hip_check(hipMemcpyAsync(ws.thl.p, ws.h_thl.data(), p.Tp * 8, hipMemcpyHostToDevice, stream),
                  "upload time grid");
    }
    ws.rtab.reserve((size_t)B * p.n_layers * p.Tp * 8);
    ws.lc.reserve((size_t)B * p.n_layers * sizeof(ekgk::LayerConst));
    if (p.mfma) {
        ws.G.reserve((size_t)B * (V + ekgk::MFMA_PAD) * 16 * 8);
        ws.aux.reserve((size_t)B * (V + ekgk::MFMA_PAD) * sizeof(float4));
    } else {
        ws.G.reserve((size_t)B * (V + 1) * (p.LP + 2) * 8);
    }
    ws.gsum.reserve((size_t)B * std::max(nch, 1) * 3 * p.LP * 8);
    ws.partial.reserve((size_t)B * std::max(nch, 1) * p.LP * p.Tp * 8);

    ekgk::LeadFieldArgs la;
    la.grid_layer = h.d_layer.as<uint8_t>();
    la.n0 = h.n0;
    la.n1 = h.n1;
    la.n2 = h.n2;
    la.bricks = Lo.bricks.as<int4>();
    la.brick_off = Lo.brick_off.as<int>();
    la.brick_vox = Lo.brick_vox.as<int2>();
    la.halo_off = Lo.halo_off.as<int>();
    la.halo_cells = Lo.halo_cells.as<uint16_t>();
    la.n_bricks = Lo.n_bricks;
    la.vox_chunk = Lo.vox_chunk.as<int>();
    la.vox_a64 = Lo.vox_a64.as<double>();
    la.vox_a = Lo.vox_a.as<float2>();
    la.chunks = Lo.chunks.as<int4>();
    la.n_chunks = nch;
    la.nhood = p.nhood;
    la.lead_pos = d_lead_pos;
    la.n_leads = p.L;
    la.layer_k = d_layer_k;
    la.n_layers = p.n_layers;
    la.V = V;
    la.LP = p.LP;
    la.n_req = B;
    la.rec = ws.G.as<double>();
    la.aux = ws.aux.as<float4>();
    la.gsum = ws.gsum.as<double>();
    la.mfma = p.mfma;
    hip_check(ekgk::launch_lead_field(la, stream), "lead_field launch");

    hip_check(ekgk::launch_ap_tables(d_layer_k, p.n_layers, ws.tgrid.as<double>(),
                                     (int)p.Tp, B, ws.rtab.as<double>(),
                                     ws.lc.as<ekgk::LayerConst>(), stream),
              "ap_tables launch");

    ekgk::FusedArgs fa;
    fa.vox_a = Lo.vox_a.as<float2>();
    fa.rec = ws.G.as<double>();
    fa.aux = ws.aux.as<float4>();
    fa.mfma = p.mfma;
    fa.n_leads = p.L;
    fa.gsum = ws.gsum.as<double>();
    fa.chunks = Lo.chunks.as<int4>();
    fa.vox_a64 = Lo.vox_a64.as<double>();
    fa.lc = ws.lc.as<ekgk::LayerConst>();
    fa.layer_k = d_layer_k;
    fa.rtab = ws.rtab.as<double>();
    fa.tgrid = ws.tgrid.as<double>();
    fa.thl = ws.thl.as<float2>();
    fa.partial = ws.partial.as<double>();
    fa.V = V;
    fa.n_layers = p.n_layers;
    fa.n_chunks = nch;
    fa.Tp = (int)p.Tp;
    fa.n_tiles = p.n_tiles;
    fa.n_groups = p.n_groups;
    fa.LP = p.LP;
    fa.n_req = B;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (time_fused) {
        e0 = ctx.get_event();
        e1 = ctx.get_event();
        hip_check(hipEventRecord(e0, stream), "hipEventRecord");
    }
    hip_check(ekgk::launch_ecg_fused(fa, stream), "ecg_fused launch");
    if (time_fused) {
        hip_check(hipEventRecord(e1, stream), "hipEventRecord");
        ctx.pending.emplace_back(e0, e1);
    }
    hip_check(ekgk::launch_ecg_reduce(ws.partial.as<double>(), d_ecg, B, nch, p.LP, p.L, (int)p.T,
                                      (int)p.Tn, (int)p.Tp, stream),
              "ecg_reduce launch");
}

void simulate(Heart& h, int batch, const ekg_sim_desc* descs, double* ecg)
{
    Ctx& ctx = *h.ctx;
    std::lock_guard<std::mutex> g1(ctx.mu);
    std::lock_guard<std::mutex> g2(h.mu);
    if (!ecg) fail(EKG_EINVAL, "ecg output buffer is required");
    BatchPlan p = plan_batch(h, batch, descs);
    if (p.T == 0) return;
    p.mfma = use_mfma(ctx, p.L);
    ctx.bind();
    ensure_layout(h);
    const size_t n_out = (size_t)p.B * p.L * p.T;
    if (h.layout.V == 0) {
        std::fill(ecg, ecg + n_out, 0.0);
        return;
    }
    upload_batch(p, descs, ctx.ws, ctx.stream);
    enqueue_batch(ctx, h, p, ctx.ws.layer_k.as<double>(), ctx.ws.lead_pos.as<double>(),
                  ctx.ws.ecg.as<double>(), ctx.ws, ctx.stream, ctx.profiling);
    hip_check(hipMemcpyAsync(ecg, ctx.ws.ecg.p, n_out * 8, hipMemcpyDeviceToHost, ctx.stream),
              "download ECG");
    hip_check(hipStreamSynchronize(ctx.stream), "hipStreamSynchronize");
    if (ctx.profiling) ctx.collect_timings();
}

void simulate_device(Heart& h, const ekg_batch_dev& d)
{
    Ctx& ctx = *h.ctx;
    std::lock_guard<std::mutex> g1(ctx.mu);
    std::lock_guard<std::mutex> g2(h.mu);
    if (!d.layer_k_dev || !d.lead_pos_dev || !d.ecg_dev)
        fail(EKG_EINVAL, "device pointers are required");
    BatchPlan p;
    p.B = d.batch;
    p.L = d.n_leads;
    p.n_layers = d.n_layers;
    p.nhood = d.nhood;
    p.start = d.start;
    p.length = d.length;
    p.dt = d.dt;
    p = plan_shape(h, p);
    if (p.T == 0) return;
    p.mfma = use_mfma(ctx, p.L);
    ctx.bind();
    ensure_layout(h);
    if (h.layout.V == 0) {
        hip_check(hipMemsetAsync(d.ecg_dev, 0, (size_t)p.B * p.L * p.T * 8, ctx.stream),
                  "hipMemsetAsync");
        return;
    }
    enqueue_batch(ctx, h, p, d.layer_k_dev, d.lead_pos_dev, d.ecg_dev, ctx.ws, ctx.stream,
                  ctx.profiling);
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
    DevBuf d_dist, d_wtab, d_delay, d_flags, d_any;
    d_dist.reserve((size_t)h.nvox * (is_int ? 4 : 8));
    d_flags.reserve((size_t)nbricks * 2);
    d_any.reserve(sizeof(int));
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
    int any = 1;
    int64_t sweeps = 0;
    const int64_t max_sweeps = 64 + 8 * (int64_t)h.nvox;
    while (any) {
        hip_check(hipMemsetAsync(fb, 0, nbricks, s), "hipMemsetAsync");
        hip_check(hipMemsetAsync(d_any.p, 0, sizeof(int), s), "hipMemsetAsync");
        hip_check(ekgk::launch_ca_sweep(is_int, h.d_layer.as<uint8_t>(), d_delay.as<uint8_t>(),
                                        d_wtab.as<double>(), cols, d_dist.p, h.n0, h.n1, h.n2, nb0,
                                        nb1, nb2, fa, fb, d_any.as<int>(), s),
                  "ca_sweep launch");
        hip_check(hipMemcpyAsync(&any, d_any.p, sizeof(int), hipMemcpyDeviceToHost, s), "read flag");
        hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
        std::swap(fa, fb);
        if (++sweeps > max_sweeps) fail(EKG_EDEVICE, "excitation relaxation did not converge");
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
