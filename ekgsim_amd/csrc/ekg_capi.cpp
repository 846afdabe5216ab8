// ekg_capi.cpp -- extern "C" entry points of include/ekg.h.  Each one
// converts library exceptions into an error code and a per-thread message;
// nothing throws across the ABI.
#include <cmath>
#include <cstring>
#include <new>
#include <string>

#include "ekg.h"
#include "ekg_engine.hpp"
#include "ekg_model.hpp"
#include "ekg_server.hpp"

struct ekg_ctx {
    ekg::Ctx c;
};
struct ekg_heart {
    ekg::Heart h;
};

namespace {

thread_local std::string g_err;

template <class F>
int guard(F&& f)
{
    try {
        f();
        g_err.clear();
        return EKG_OK;
    } catch (const ekg::Error& e) {
        g_err = e.msg;
        return e.code;
    } catch (const std::bad_alloc&) {
        g_err = "host allocation failed";
        return EKG_ENOMEM;
    } catch (const std::exception& e) {
        g_err = e.what();
        return EKG_EDEVICE;
    } catch (...) {
        g_err = "unknown failure";
        return EKG_EDEVICE;
    }
}

// U/V basis of sim_lib.h:181-193 in the reference's (z,y,x) storage:
// x-hat = (0,0,1); U = -(x-hat x p), V = (x-hat x p) x p, both normalised
// by multiplying with 1/|.| (Hypermatrix.h:264-270; 0 stays 0).
void basis(const double p[3], double u[3], double v[3])
{
    const double x[3] = {0, 0, 1};
    double a[3], b[3];
    a[0] = x[1] * p[2] - x[2] * p[1];
    a[1] = x[2] * p[0] - x[0] * p[2];
    a[2] = x[0] * p[1] - x[1] * p[0];
    b[0] = a[1] * p[2] - a[2] * p[1];
    b[1] = a[2] * p[0] - a[0] * p[2];
    b[2] = a[0] * p[1] - a[1] * p[0];
    a[0] *= -1;
    a[1] *= -1;
    a[2] *= -1;
    for (double* w : {a, b}) {
        double l = w[0] * w[0];
        l += w[1] * w[1];
        l += w[2] * w[2];
        l = std::sqrt(l);
        const double s = (l != 0) ? 1.0 / l : 0.0;
        w[0] *= s;
        w[1] *= s;
        w[2] *= s;
    }
    std::memcpy(u, a, sizeof a);
    std::memcpy(v, b, sizeof b);
}

}  // namespace

extern "C" {

const char* ekg_last_error(void) { return g_err.c_str(); }

const char* ekg_version(void)
{
    return "ekgsim-amd 0.1 (gfx950; FP32 AP, FP64 lead sums; bit-exact FP64/int32 excitation)";
}

int ekg_ctx_create(int device, ekg_ctx** out)
{
    return guard([&] {
        if (!out) ekg::fail(EKG_EINVAL, "out is NULL");
        *out = nullptr;
        int n = 0;
        ekg::hip_check(hipGetDeviceCount(&n), "hipGetDeviceCount");
        if (device < 0 || device >= n)
            ekg::fail(EKG_EINVAL, "device " + std::to_string(device) + " not present (" +
                                      std::to_string(n) + " visible)");
        ekg_ctx* c = new ekg_ctx;
        try {
            c->c.device = device;
            c->c.bind();
            ekg::hip_check(hipStreamCreateWithFlags(&c->c.stream, hipStreamNonBlocking),
                           "hipStreamCreate");
            // the device's static workspace (allocation-free entry point),
            // resolved once per device with that device current
            void* scratch = nullptr;
            ekg::hip_check(ekgk::scratch_ptr(device, &scratch), "resolve the static device workspace");
            c->c.ws_fixed.arena = static_cast<char*>(scratch);
            c->c.ws_fixed.arena_bytes = ekgk::SCRATCH_BYTES;
            // second streams of the launch groups, made now so that no later
            // call (the allocation-free one included) creates them
            c->c.ws.ensure_side();
            c->c.ws_fixed.ensure_side();
            // per neighbourhood: {d0, d1, d2, index of d among the 26 cube offsets}
            int8_t tab[4][26][4] = {};
            int8_t cube[26][3];
            ekg::nhood_offsets(EKG_NHOOD_3D8, cube);
            for (int t = 0; t < 4; ++t) {
                int8_t d[26][3];
                const int n = ekg::nhood_offsets(t, d);
                for (int q = 0; q < n; ++q) {
                    int bit = 0;
                    while (cube[bit][0] != d[q][0] || cube[bit][1] != d[q][1] || cube[bit][2] != d[q][2])
                        ++bit;
                    tab[t][q][0] = d[q][0];
                    tab[t][q][1] = d[q][1];
                    tab[t][q][2] = d[q][2];
                    tab[t][q][3] = (int8_t)bit;
                }
            }
            c->c.nb_table.reserve(sizeof tab);
            ekg::hip_check(hipMemcpy(c->c.nb_table.p, tab, sizeof tab, hipMemcpyHostToDevice),
                           "upload neighbourhoods");
        } catch (...) {
            if (c->c.stream) (void)hipStreamDestroy(c->c.stream);
            delete c;
            throw;
        }
        *out = c;
    });
}

void ekg_ctx_destroy(ekg_ctx* ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->c.device);
    (void)hipStreamSynchronize(ctx->c.stream);
    for (auto& pr : ctx->c.pending) {
        (void)hipEventDestroy(pr.first);
        (void)hipEventDestroy(pr.second);
    }
    for (auto e : ctx->c.free_events) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(ctx->c.stream);
    delete ctx;
}

int ekg_ctx_set_profiling(ekg_ctx* ctx, int enable)
{
    return guard([&] {
        if (!ctx) ekg::fail(EKG_EINVAL, "ctx is NULL");
        std::lock_guard<std::mutex> g(ctx->c.mu);
        ctx->c.profiling = enable != 0;
    });
}

int ekg_ctx_kernel_time(ekg_ctx* ctx, const char* kernel, int64_t* launches, double* total_ms)
{
    return guard([&] {
        if (!ctx || !launches || !total_ms) ekg::fail(EKG_EINVAL, "NULL argument");
        if (kernel && std::strcmp(kernel, "ecg_fused") != 0)
            ekg::fail(EKG_EINVAL, "only the ecg_fused kernel is timed");
        std::lock_guard<std::mutex> g(ctx->c.mu);
        ctx->c.bind();
        ctx->c.collect_timings();
        *launches = ctx->c.fused_launches;
        *total_ms = ctx->c.fused_ms;
    });
}

int ekg_ctx_set_kernel(ekg_ctx* ctx, int kind)
{
    return guard([&] {
        if (!ctx) ekg::fail(EKG_EINVAL, "ctx is NULL");
        if (kind < EKG_KERNEL_AUTO || kind > EKG_KERNEL_MFMA) ekg::fail(EKG_EINVAL, "unknown kernel kind");
        std::lock_guard<std::mutex> g(ctx->c.mu);
        ctx->c.kernel = kind;
    });
}

int ekg_ctx_reset_profiling(ekg_ctx* ctx)
{
    return guard([&] {
        if (!ctx) ekg::fail(EKG_EINVAL, "ctx is NULL");
        std::lock_guard<std::mutex> g(ctx->c.mu);
        ctx->c.bind();
        ctx->c.collect_timings();
        ctx->c.fused_launches = 0;
        ctx->c.fused_ms = 0.0;
    });
}

int ekg_heart_create(ekg_ctx* ctx, int n0, int n1, int n2, const uint8_t* layer,
                     const uint8_t* seed, ekg_heart** out)
{
    return guard([&] {
        if (!ctx || !layer || !out) ekg::fail(EKG_EINVAL, "NULL argument");
        *out = nullptr;
        if (n0 < 1 || n1 < 1 || n2 < 1) ekg::fail(EKG_EINVAL, "grid dimensions must be >= 1");
        const int64_t nvox = (int64_t)n0 * n1 * n2;
        if (nvox > 0x7fffffffLL) ekg::fail(EKG_EINVAL, "grid larger than 2^31 voxels");
        std::lock_guard<std::mutex> g(ctx->c.mu);
        ekg_heart* H = new ekg_heart;
        try {
            ekg::Heart& h = H->h;
            h.ctx = &ctx->c;
            h.device = ctx->c.device;
            h.n0 = n0;
            h.n1 = n1;
            h.n2 = n2;
            h.nvox = nvox;
            h.h_layer.assign(layer, layer + nvox);
            if (seed) h.h_seed.assign(seed, seed + nvox);
            for (int64_t i = 0; i < nvox; ++i) {
                if (h.h_layer[i]) ++h.n_myo;
                if (h.h_layer[i] > h.max_layer) h.max_layer = h.h_layer[i];
            }
            ctx->c.bind();
            h.d_layer.reserve(nvox);
            ekg::hip_check(hipMemcpy(h.d_layer.p, layer, nvox, hipMemcpyHostToDevice), "upload layers");
            h.d_seed.reserve(nvox);
            if (seed)
                ekg::hip_check(hipMemcpy(h.d_seed.p, seed, nvox, hipMemcpyHostToDevice), "upload seeds");
            else
                ekg::hip_check(hipMemset(h.d_seed.p, 0, nvox), "clear seeds");
        } catch (...) {
            delete H;
            throw;
        }
        *out = H;
    });
}

void ekg_heart_destroy(ekg_heart* h)
{
    if (!h) return;
    (void)hipSetDevice(h->h.device);
    delete h;
}

int ekg_heart_num_myocardial(const ekg_heart* h, int64_t* n)
{
    return guard([&] {
        if (!h || !n) ekg::fail(EKG_EINVAL, "NULL argument");
        *n = h->h.n_myo;
    });
}

int ekg_heart_max_layer(const ekg_heart* h, int* max_layer)
{
    return guard([&] {
        if (!h || !max_layer) ekg::fail(EKG_EINVAL, "NULL argument");
        *max_layer = h->h.max_layer;
    });
}

int ekg_heart_excite_transfer(ekg_heart* h, const double* transfer, int rows, int cols)
{
    return guard([&] {
        if (!h) ekg::fail(EKG_EINVAL, "heart is NULL");
        ekg::excite(h->h, false, transfer, rows, cols, nullptr);
    });
}

int ekg_heart_excite_delay(ekg_heart* h, const uint8_t* delay)
{
    return guard([&] {
        if (!h) ekg::fail(EKG_EINVAL, "heart is NULL");
        ekg::excite(h->h, true, nullptr, 0, 0, delay);
    });
}

int ekg_heart_set_activation(ekg_heart* h, const double* act)
{
    return guard([&] {
        if (!h || !act) ekg::fail(EKG_EINVAL, "NULL argument");
        ekg::Heart& H = h->h;
        std::lock_guard<std::mutex> g1(H.ctx->mu);
        std::lock_guard<std::mutex> g2(H.mu);
        for (int64_t i = 0; i < H.nvox; ++i)
            if (!std::isfinite(act[i])) ekg::fail(EKG_EINVAL, "non-finite activation time");
        H.h_act.assign(act, act + H.nvox);
        H.ctx->bind();
        H.d_act.reserve((size_t)H.nvox * 8);
        ekg::hip_check(hipMemcpy(H.d_act.p, act, (size_t)H.nvox * 8, hipMemcpyHostToDevice),
                       "upload activation");
        H.excited = true;
        H.layout.valid = false;
    });
}

int ekg_heart_get_activation(const ekg_heart* h, double* act)
{
    return guard([&] {
        if (!h || !act) ekg::fail(EKG_EINVAL, "NULL argument");
        if (!h->h.excited) ekg::fail(EKG_ESTATE, "excitation sequence not computed or loaded");
        std::memcpy(act, h->h.h_act.data(), (size_t)h->h.nvox * 8);
    });
}

int64_t ekg_num_steps(double length, double dt) { return ekg::num_steps(length, dt); }

int ekg_simulate(ekg_heart* h, const ekg_sim_desc* d, double* ecg)
{
    return guard([&] {
        if (!h || !d) ekg::fail(EKG_EINVAL, "NULL argument");
        ekg::simulate(h->h, 1, d, ecg);
    });
}

int ekg_simulate_batch(ekg_heart* h, int batch, const ekg_sim_desc* descs, double* ecg)
{
    return guard([&] {
        if (!h || !descs) ekg::fail(EKG_EINVAL, "NULL argument");
        ekg::simulate(h->h, batch, descs, ecg);
    });
}

int ekg_simulate_batch_device(ekg_heart* h, const ekg_batch_dev* d)
{
    return guard([&] {
        if (!h || !d) ekg::fail(EKG_EINVAL, "NULL argument");
        ekg::simulate_device(h->h, *d);
    });
}

int ekg_simulate_batch_device_noalloc(ekg_heart* h, const ekg_batch_dev* d)
{
    return guard([&] {
        if (!h || !d) ekg::fail(EKG_EINVAL, "NULL argument");
        ekg::simulate_device_fixed(h->h, *d);
    });
}

int ekg_ctx_synchronize(ekg_ctx* ctx)
{
    return guard([&] {
        if (!ctx) ekg::fail(EKG_EINVAL, "ctx is NULL");
        std::lock_guard<std::mutex> g(ctx->c.mu);
        ctx->c.bind();
        ekg::hip_check(hipStreamSynchronize(ctx->c.stream), "hipStreamSynchronize");
    });
}

void ekg_lead_basis(const double p[3], double u[3], double v[3]) { basis(p, u, v); }

void ekg_move_lead(const double p[3], double du, double dv, double out[3])
{
    double u[3], v[3];
    basis(p, u, v);
    for (int k = 0; k < 3; ++k) out[k] = (p[k] + u[k] * du) + v[k] * dv;
}

// ---- tuning and graphs ------------------------------------------------------

int ekg_heart_autotune(ekg_heart* h, const ekg_sim_desc* sample, int batch, ekg_tune* out)
{
    return guard([&] {
        if (!h || !sample) ekg::fail(EKG_EINVAL, "NULL argument");
        const ekg_tune t = ekg::autotune(h->h, *sample, batch);
        if (out) *out = t;
    });
}

int ekg_heart_set_tune(ekg_heart* h, const ekg_tune* t)
{
    return guard([&] {
        if (!h || !t) ekg::fail(EKG_EINVAL, "NULL argument");
        if (t->chunk_voxels != 0 && (t->chunk_voxels < 64 || t->chunk_voxels > ekgk::CHUNK_MAX))
            ekg::fail(EKG_EINVAL, "chunk_voxels must be 0 or 64..512");
        if (t->mfma_waves != 0 && (t->mfma_waves < 2 || t->mfma_waves > 16))
            ekg::fail(EKG_EINVAL, "mfma_waves must be 0 or 2..16");
        if (t->lead_phase != 0 && t->lead_phase != 2 && t->lead_phase != 4)
            ekg::fail(EKG_EINVAL, "lead_phase must be 0, 2 or 4");
        if (t->kernel < EKG_KERNEL_AUTO || t->kernel > EKG_KERNEL_MFMA) ekg::fail(EKG_EINVAL, "unknown kernel");
        if (t->class_min < -1) ekg::fail(EKG_EINVAL, "class_min must be -1, 0 or a voxel count");
        ekg::Ctx& c = *h->h.ctx;
        std::lock_guard<std::mutex> g1(c.mu);
        std::lock_guard<std::mutex> g2(h->h.mu);
        h->h.tune_chunk = t->chunk_voxels;
        h->h.tune_waves = t->mfma_waves;
        h->h.tune_phase = t->lead_phase;
        h->h.tune_kernel = t->kernel;
        h->h.tune_class = t->class_min;
    });
}

int ekg_heart_get_tune(const ekg_heart* h, ekg_tune* t)
{
    return guard([&] {
        if (!h || !t) ekg::fail(EKG_EINVAL, "NULL argument");
        t->chunk_voxels = h->h.tune_chunk;
        t->mfma_waves = h->h.tune_waves;
        t->lead_phase = h->h.tune_phase;
        t->kernel = h->h.tune_kernel;
        t->class_min = h->h.tune_class;
        t->ms_per_batch = 0.0;
    });
}

int ekg_criteria_device(ekg_ctx* ctx, const ekg_criteria_desc* c, const double* d_ecg, int batch, int n_leads,
                        int64_t n_steps, double* d_out)
{
    return guard([&] {
        if (!ctx || !c) ekg::fail(EKG_EINVAL, "NULL argument");
        ekg::criteria_device(ctx->c, *c, d_ecg, batch, n_leads, n_steps, d_out);
    });
}

int ekg_simulate_criteria(ekg_heart* h, int batch, const ekg_sim_desc* descs, const ekg_criteria_desc* c,
                          double* criteria)
{
    return guard([&] {
        if (!h || !c) ekg::fail(EKG_EINVAL, "NULL argument");
        ekg::simulate_criteria(h->h, batch, descs, *c, criteria);
    });
}

int ekg_heart_layout_info(ekg_heart* h, int nhood, ekg_layout_info* out)
{
    return guard([&] {
        if (!h || !out) ekg::fail(EKG_EINVAL, "NULL argument");
        if (nhood < EKG_NHOOD_2D4 || nhood > EKG_NHOOD_3D8) ekg::fail(EKG_EINVAL, "unknown neighbourhood");
        ekg::Ctx& c = *h->h.ctx;
        std::lock_guard<std::mutex> g1(c.mu);
        std::lock_guard<std::mutex> g2(h->h.mu);
        ekg::ensure_layout(h->h, true);
        ekg::ensure_cells(h->h, nhood);
        const ekg::EcgLayout& L = h->h.layout;
        ekg_layout_info o{};
        o.voxels = L.V;
        for (const int4& cc : L.h_cchunks) o.class_voxels += cc.z - cc.y;
        o.class_cells = L.n_ccells[nhood];
        o.chunks = L.n_chunks;
        o.class_chunks = L.n_cchunks;
        o.classes = L.n_classes;
        o.class_pieces = L.n_cpieces[nhood];
        o.class_entries = L.n_centries[nhood];
        *out = o;
    });
}

int ekg_ctx_set_graphs(ekg_ctx* ctx, int enable)
{
    return guard([&] {
        if (!ctx) ekg::fail(EKG_EINVAL, "ctx is NULL");
        std::lock_guard<std::mutex> g(ctx->c.mu);
        ctx->c.graphs = enable != 0;
    });
}

int ekg_ctx_graph_stats(ekg_ctx* ctx, int64_t* captures, int64_t* launches)
{
    return guard([&] {
        if (!ctx) ekg::fail(EKG_EINVAL, "ctx is NULL");
        std::lock_guard<std::mutex> g(ctx->c.mu);
        if (captures) *captures = ctx->c.graph_captures;
        if (launches) *launches = ctx->c.graph_launches;
    });
}

// ---- parameter vector -> layer APs, comparison ----------------------------

double ekg_ap_value(const double k[9], double t) { return ekg::ap_value(k, t); }

double ekg_ap_apd90(const double k[9]) { return ekg::ap_apd90(k); }

int ekg_ap_num_genes(const ekg_ap_model* m, int* n_genes)
{
    return guard([&] {
        if (!m || !n_genes) ekg::fail(EKG_EINVAL, "null argument");
        if (m->interpolation != EKG_INTERP_ENDO_EPI && m->interpolation != EKG_INTERP_ENDO_MID_EPI)
            ekg::fail(EKG_EINVAL, "unknown interpolation type");
        if (m->n_free < 0 || (m->n_free > 0 && !m->free_k)) ekg::fail(EKG_EINVAL, "bad free-k list");
        *n_genes = ekg::ap_num_genes(*m);
    });
}

int ekg_layer_aps(const ekg_ap_model* m, int n_layers, const double* genes, int n_genes, double* layer_k,
                  double* violation)
{
    return guard([&] {
        if (!m) ekg::fail(EKG_EINVAL, "null model");
        ekg::layer_aps(*m, n_layers, genes, n_genes, layer_k, violation);
    });
}

int ekg_resample(const double* src, int64_t n, double factor, int start, double* dest, int64_t cap,
                 int64_t* n_out)
{
    return guard([&] {
        if (!src || n <= 0 || !n_out || cap < 0 || (cap > 0 && !dest)) ekg::fail(EKG_EINVAL, "bad resample arguments");
        *n_out = (int64_t)ekg::resample(src, (size_t)n, factor, start, dest, (size_t)cap);
    });
}

int ekg_criterion(int mode, const double* sim, int64_t n_sim, const double* target, int64_t n_target,
                  double target_offset, double* value)
{
    return guard([&] {
        if (!value || n_sim < 0 || n_target < 0 || (n_sim > 0 && !sim) || (n_target > 0 && !target))
            ekg::fail(EKG_EINVAL, "bad criterion arguments");
        *value = ekg::criterion(mode, sim, (size_t)n_sim, target, (size_t)n_target, target_offset);
    });
}

// ---- server ---------------------------------------------------------------

int ekg_server_create(ekg_ctx* ctx, const ekg_server_opts* opts, ekg_server** out)
{
    return guard([&] {
        if (!ctx || !out) ekg::fail(EKG_EINVAL, "NULL argument");
        *out = nullptr;
        ekg_server_opts o{};
        if (opts) o = *opts;
        *out = reinterpret_cast<ekg_server*>(ekg::server_create(ctx->c, o));
    });
}

void ekg_server_destroy(ekg_server* s)
{
    if (s) ekg::server_destroy(reinterpret_cast<ekg::Server*>(s));
}

int ekg_server_add_heart(ekg_server* s, ekg_heart* h, int* heart_id)
{
    return guard([&] {
        if (!s || !h || !heart_id) ekg::fail(EKG_EINVAL, "NULL argument");
        *heart_id = ekg::server_add_heart(*reinterpret_cast<ekg::Server*>(s), h->h);
    });
}

int ekg_server_submit(ekg_server* s, int heart_id, const ekg_sim_desc* d, uint64_t* ticket)
{
    return guard([&] {
        if (!s || !d || !ticket) ekg::fail(EKG_EINVAL, "NULL argument");
        *ticket = ekg::server_submit(*reinterpret_cast<ekg::Server*>(s), heart_id, *d);
    });
}

int ekg_server_poll(ekg_server* s, uint64_t ticket, int* done)
{
    return guard([&] {
        if (!s || !done) ekg::fail(EKG_EINVAL, "NULL argument");
        *done = ekg::server_poll(*reinterpret_cast<ekg::Server*>(s), ticket) ? 1 : 0;
    });
}

int ekg_server_wait(ekg_server* s, uint64_t ticket, double* ecg, size_t cap, double* latency_ms)
{
    return guard([&] {
        if (!s) ekg::fail(EKG_EINVAL, "NULL argument");
        ekg::server_wait(*reinterpret_cast<ekg::Server*>(s), ticket, ecg, cap, latency_ms);
    });
}

int ekg_server_wait_view(ekg_server* s, uint64_t ticket, const double** ecg, size_t* n, double* latency_ms)
{
    return guard([&] {
        if (!s || !ecg) ekg::fail(EKG_EINVAL, "NULL argument");
        *ecg = ekg::server_wait_view(*reinterpret_cast<ekg::Server*>(s), ticket, n, latency_ms);
    });
}

int ekg_server_release(ekg_server* s, uint64_t ticket)
{
    return guard([&] {
        if (!s) ekg::fail(EKG_EINVAL, "NULL argument");
        ekg::server_release(*reinterpret_cast<ekg::Server*>(s), ticket);
    });
}

int ekg_server_stats(ekg_server* s, uint64_t* completed, uint64_t* groups, int* peak_inflight,
                     size_t* peak_bytes)
{
    return guard([&] {
        if (!s) ekg::fail(EKG_EINVAL, "NULL argument");
        ekg::server_stats(*reinterpret_cast<ekg::Server*>(s), completed, groups, peak_inflight,
                          peak_bytes);
    });
}

}  // extern "C"
