/*
 * ekg.h -- C ABI of the MI355X-native ekgsim simulation core.
 *
 * The reference (synergy-twinning/ekgsim) exposes the simulation core as the
 * C++ class SimLib::EkgSim (simlib/sim_lib.h:93-287) over SimLib::Simulation
 * (simlib/simulator.h:508-586).  This header is the thin C layer a binding
 * (ctypes, cgo, JNI, N-API) or the C++ facade in ekgsim_amd/csrc/sim_lib.hpp
 * calls.  Conventions:
 *   - every call returns EKG_OK (0) or an error code; the message of the last
 *     failure on the calling thread is ekg_last_error();
 *   - no exceptions cross the ABI; all buffers are caller-owned host memory;
 *   - voxel arrays are row-major (i0, i1, i2) = (z, y, x) with i2 fastest,
 *     the reference's storage order (copyOfLibs/Pattern/Hypermatrix.h:652-661);
 *   - lead positions are (z, y, x) in voxel units relative to the model
 *     corner, as SimLib::MeasuringPoint stores them (simulator.cpp:369-390);
 *   - an ECG is n_leads rows of T = ceil(length/dt) FP64 samples
 *     (simulator.cpp:470), lead-major.
 * A context owns one GPU and one HIP stream; calls on one context must not
 * overlap (the reference is single-threaded too, sim.cpp:444).  Servers have
 * their own internal locking.
 */
#ifndef EKG_H_INCLUDED
#define EKG_H_INCLUDED

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EKG_OK 0
#define EKG_EINVAL 1      /* bad argument or shape */
#define EKG_ENOMEM 2      /* host or device allocation failed / over budget */
#define EKG_EDEVICE 3     /* HIP runtime error */
#define EKG_ENOSEED 4     /* no excitation starting point (simulator.cpp:276-279) */
#define EKG_ETRANSFER 5   /* conduction matrix too small / negative (simulator.cpp:201-205, 233-237) */
#define EKG_ESTATE 6      /* call order violated (e.g. simulate before excitation) */
#define EKG_ENOTFOUND 7   /* unknown ticket / heart id */
#define EKG_EIO 8         /* file read/parse failure */

/* Dipole neighbourhoods, `[simulation] neighbourhood type` (sim_lib.h:131-143).
 * 3D4 is the 8 corner neighbours (+-1,+-1,+-1) (simulator.h:349-357). */
#define EKG_NHOOD_2D4 0
#define EKG_NHOOD_2D8 1
#define EKG_NHOOD_3D4 2
#define EKG_NHOOD_3D8 3

typedef struct ekg_ctx ekg_ctx;
typedef struct ekg_heart ekg_heart;
typedef struct ekg_server ekg_server;

/* Message of the last failed call on this thread ("" if none). */
const char *ekg_last_error(void);

/* Library build string (kernel target, features). */
const char *ekg_version(void);

/* ---- context ------------------------------------------------------------ */
/* One GPU, one stream.  No reference counterpart (the reference is CPU-only);
 * it replaces the process-wide state of SimLib::EkgSim. */
int ekg_ctx_create(int device, ekg_ctx **out);
/* Destroy the context's servers first.  A heart may be destroyed before or
 * after its context (it frees only its own device memory); no other call on
 * a heart is valid once its context is destroyed. */
void ekg_ctx_destroy(ekg_ctx *ctx);
/* Kernel timing on the context stream: when enabled, every launch of the
 * fused AP+lead kernel is bracketed by HIP events; ekg_ctx_kernel_time
 * returns the number of launches and their summed duration since the last
 * reset. */
int ekg_ctx_set_profiling(ekg_ctx *ctx, int enable);
int ekg_ctx_kernel_time(ekg_ctx *ctx, const char *kernel, int64_t *launches, double *total_ms);
int ekg_ctx_reset_profiling(ekg_ctx *ctx);

/* Lead-contraction path of the fused AP+lead kernel: the matrix cores
 * (v_mfma_f64_16x16x4_f64, FP64 accumulate) or the vector ALUs (FP64 FMA
 * with the lead weights in scalar registers).  Both give the same ECGs to
 * FP64 rounding; AUTO picks the matrix cores. */
#define EKG_KERNEL_AUTO 0
#define EKG_KERNEL_VALU 1
#define EKG_KERNEL_MFMA 2
int ekg_ctx_set_kernel(ekg_ctx *ctx, int kind);

/* ---- heart (device-resident geometry) ----------------------------------- */
/* Replaces EkgSim::loadShape (sim_lib.h:211-216) -> Simulation::loadShape
 * (simulator.cpp:181-199).  layer[i] = 0 empty, 1..255 myocardial layer;
 * seed[i] != 0 marks an excitation start (the 0x1000 flag of matrix.h:87,
 * a negative value in the .matrix file, matrix.h:124-129).  seed may be
 * NULL when activation times are supplied with ekg_heart_set_activation. */
int ekg_heart_create(ekg_ctx *ctx, int n0, int n1, int n2, const uint8_t *layer,
                     const uint8_t *seed, ekg_heart **out);
void ekg_heart_destroy(ekg_heart *h);
int ekg_heart_num_myocardial(const ekg_heart *h, int64_t *n);
int ekg_heart_max_layer(const ekg_heart *h, int *max_layer);

/* Excitation sequence on the GPU, FP64, bit-exact with
 * Simulation::calculateExcitationSequence (simulator.cpp:248-286) /
 * exciteElement (simulator.cpp:212-246): seeds at 1.0, lag
 * transfer[layer(u)][layer(v)] * sqrt(|d|^2) over the 26-neighbourhood (8 in
 * 2D), 0.0 for unreached myocardium.  transfer is row-major rows x cols;
 * every myocardial layer must be < rows and < cols and every used entry must
 * be >= 0 (EKG_ETRANSFER otherwise). */
int ekg_heart_excite_transfer(ekg_heart *h, const double *transfer, int rows, int cols);

/* Excitation sequence with an integer per-voxel conduction delay (the
 * synthetic-heart configuration; no reference counterpart): seeds at 1, lag
 * delay[v] steps into voxel v from any of its 26 (8 in 2D) neighbours. */
int ekg_heart_excite_delay(ekg_heart *h, const uint8_t *delay);

/* Activation times supplied by the caller, one per voxel of the grid
 * (Simulation::loadExcitationSequence, simulator.cpp:288-367). */
int ekg_heart_set_activation(ekg_heart *h, const double *act);
/* Activation times of the whole grid (excitationDelay of every
 * ShapeElement, matrix.h:84). */
int ekg_heart_get_activation(const ekg_heart *h, double *act);

/* ---- plain simulation (the drop-in) ------------------------------------- */
typedef struct ekg_sim_desc {
    int n_layers;           /* rows of layer_k; must be >= the heart's max layer */
    const double *layer_k;  /* [n_layers][9] WohlfartPlus coefficients; layer l uses row l-1
                               (Simulation::setApIndices, simulator.cpp:561-621) */
    int n_leads;            /* 1..EKG_MAX_LEADS */
    const double *lead_pos; /* [n_leads][3] (z,y,x) measuring points, already displaced */
    int nhood;              /* EKG_NHOOD_* (dipole neighbourhood) */
    double start;           /* [simulation] start, ms */
    double length;          /* [simulation] length, ms */
    double dt;              /* [simulation] time step, ms */
} ekg_sim_desc;

#define EKG_MAX_LEADS 16

/* Samples per lead: ceil(length / dt) (simulator.cpp:470). */
int64_t ekg_num_steps(double length, double dt);

/* One ECG, replacing Simulation::run (simulator.cpp:452-550) and the
 * measurements EkgSim::getMeasurement returns (sim_lib.h:259-262).  AP
 * evaluation in FP32, lead sums and the cross-block reduction in FP64.
 * ecg: [n_leads][T] FP64. */
int ekg_simulate(ekg_heart *h, const ekg_sim_desc *d, double *ecg);

/* B ECGs of one heart in one set of launches (one launch per kernel).  The
 * descriptors must share nhood; lead count, layer count, start, length and
 * time step may differ per entry.  ECG b is written at offset
 * sum_{q<b} n_leads_q * T_q, as [n_leads_b][T_b]: for equal shapes that is
 * [batch][n_leads][T].  Every ECG is bitwise the same as from ekg_simulate. */
int ekg_simulate_batch(ekg_heart *h, int batch, const ekg_sim_desc *descs, double *ecg);

/* B ECGs whose per-request arrays already live in device memory (HBM):
 * layer_k_dev [batch][n_layers][9], lead_pos_dev [batch][n_leads][3], output
 * ecg_dev [batch][n_leads][T], all FP64.  Same computation as
 * ekg_simulate_batch (bitwise equal results); the work is enqueued on the
 * context stream and the call returns without waiting (ekg_ctx_synchronize
 * waits).  Values are not inspected on the host: non-finite inputs give
 * non-finite ECGs.  No reference counterpart (the reference has no device);
 * it is the form a device-resident caller such as a batched optimiser uses. */
typedef struct ekg_batch_dev {
    int batch;
    int n_layers;
    const double *layer_k_dev;
    int n_leads;
    const double *lead_pos_dev;
    int nhood;
    double start, length, dt;
    double *ecg_dev;
} ekg_batch_dev;

int ekg_simulate_batch_device(ekg_heart *h, const ekg_batch_dev *d);

/* ekg_simulate_batch_device without any device allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same kernels and bitwise the same ECGs; every group buffer
 * (tables, time grids, AP tables, chunk and class sums, prefix tables,
 * partial rows) lives in a 256 MiB static device workspace that each GPU holds
 * from library load until the process ends.  Limits and conditions:
 *   - the heart's layout for d->nhood and its current tuning must already be
 *     built, by any earlier simulation of the heart on the same contraction
 *     path or, for the matrix-core path, by ekg_heart_layout_info; else
 *     EKG_ESTATE;
 *   - a batch whose buffers exceed the workspace fails with EKG_ENOMEM before
 *     any kernel runs;
 *   - launches are eager (no graph capture of this library's own), and no
 *     kernel timing is taken;
 *   - every call on one GPU shares that GPU's workspace: calls that overlap on
 *     one GPU (two contexts, two threads, or this call next to a server) race
 *     and must be serialised by the caller.  Nothing needs to start at zero:
 *     each call writes every workspace byte it reads. */
int ekg_simulate_batch_device_noalloc(ekg_heart *h, const ekg_batch_dev *d);
int ekg_ctx_synchronize(ekg_ctx *ctx);

/* ---- launch tuning and graph replay (no reference counterpart) ---------- */
/* Tunable launch shape of a heart's simulations.  Every setting gives ECGs
 * within the same tolerance; they differ in speed only. */
typedef struct ekg_tune {
    int chunk_voxels;     /* target myocardial voxels per chunk, 64..512 */
    int mfma_waves;       /* waves per block of the matrix-core chunk kernel, 2..16 */
    int lead_phase;       /* leads per lead-field phase, 2 or 4 */
    int kernel;           /* EKG_KERNEL_MFMA or EKG_KERNEL_VALU */
    int class_min;        /* activation classes on the matrix-core path: a run of at least
                             this many voxels of one layer with one activation time is
                             summed as one class (all its voxels share the AP); 0 = the
                             default (64), -1 = no classes (per-voxel APs throughout) */
    double ms_per_batch;  /* measured device time of the chosen shape (autotune only) */
} ekg_tune;

/* Autotuner: times the launch-shape knobs on `batch` copies of `sample` on
 * this heart's GPU -- the contraction path, then every (chunk size, waves per
 * block) pair of the matrix-core path, or the chunk size and lead-field phase
 * width of the vector-ALU path -- and keeps the fastest for the heart's later
 * simulations; `out` (optional) receives the choice.  An explicit
 * ekg_ctx_set_kernel choice still overrides the tuned path. */
int ekg_heart_autotune(ekg_heart *h, const ekg_sim_desc *sample, int batch, ekg_tune *out);
/* Apply a stored tuning (e.g. cached from an earlier autotune on the same
 * heart and shape); chunk_voxels 0 restores the default layout. */
int ekg_heart_set_tune(ekg_heart *h, const ekg_tune *t);
int ekg_heart_get_tune(const ekg_heart *h, ekg_tune *t);

/* Layout of the heart's matrix-core path as currently tuned (built on
 * demand, so that later calls find it ready): chunks, activation classes and,
 * for dipole neighbourhood `nhood`, the class cells and the class-window
 * entries whose lead weights each simulation evaluates once.  No counterpart
 * in the reference (diagnostics of this library's layout). */
typedef struct ekg_layout_info {
    int64_t voxels;        /* myocardial voxels */
    int64_t class_voxels;  /* of which in activation classes */
    int64_t class_cells;   /* (class, cell) pairs with a non-zero dipole vector */
    int chunks;            /* chunks of the voxels outside classes */
    int class_chunks;
    int classes;
    int class_pieces;      /* class-window pieces (blocks per request of the class-window kernel) */
    int64_t class_entries; /* (class window, cell) pairs: lead weights evaluated per lead and request */
} ekg_layout_info;
int ekg_heart_layout_info(ekg_heart *h, int nhood, ekg_layout_info *out);

/* Graph mode: the kernel launches of a simulation group are captured
 * into a hipGraph on first use and replayed while the group keeps its
 * shape and buffers (a repeated batch with new coefficients or lead
 * positions in the same device arrays replays without recapture).  Off by
 * default; groups timed with ekg_ctx_set_profiling run eagerly. */
int ekg_ctx_set_graphs(ekg_ctx *ctx, int enable);
int ekg_ctx_graph_stats(ekg_ctx *ctx, int64_t *captures, int64_t *launches);

/* ---- lead geometry helpers (host, FP64) --------------------------------- */
/* U/V displacement basis of a measuring point (EkgSim::loadMeasuringPoints,
 * sim_lib.h:167-193) and the displaced position p + du*U + dv*V
 * (EkgSim::moveMeasuringPoints, sim_lib.h:197-208). */
void ekg_lead_basis(const double p[3], double u[3], double v[3]);
void ekg_move_lead(const double p[3], double du, double dv, double out[3]);

/* ---- parameter vector -> layer APs (host, FP64) ------------------------- */
/* The `[wohlfart ap]` section of simulator.ini (SimSettings.h:176-237). */
#define EKG_INTERP_ENDO_EPI 1     /* `interpolation = endo-epi` */
#define EKG_INTERP_ENDO_MID_EPI 2 /* `interpolation = endo-mid-epi` */

typedef struct ekg_ap_model {
    int n_base;             /* rows of base_k, `base ap 1..n` */
    const double *base_k;   /* [n_base][9] */
    const double *k_min;    /* [9] `k min` */
    const double *k_max;    /* [9] `k max` */
    int n_free;             /* entries of free_k */
    const int *free_k;      /* `free k`: 0-based coefficient indices */
    int interpolation;      /* EKG_INTERP_* */
    double mid_position;    /* `mid AP position`, in [0, 1] */
} ekg_ap_model;

/* WohlfartPlus W(t) for coefficients k[0..8] (WohlfartPlus::operator[],
 * Wohlfart.h:195-203), FP64 in the reference's operation order. */
double ekg_ap_value(const double k[9], double t);

/* APD90 of a WohlfartPlus curve (WohlfartPlus::apd90, Wohlfart.h:206-223):
 * the last time on the 1 ms grid 0..999 at which the curve falls through
 * k0 + k2/10, placed between the two samples by linear interpolation; -1 when
 * it never does. */
double ekg_ap_apd90(const double k[9]);

/* Number of parameter-vector entries the layer APs take: distinct free
 * coefficients x 2 (endo-epi) or x 3 (endo-mid-epi) (sim.cpp:492-509). */
int ekg_ap_num_genes(const ekg_ap_model *m, int *n_genes);

/* Layer APs of an n_layers heart from the leading entries of a parameter
 * vector (SimImplementation::simUsingBorderAps / simUsingBorderAndMidAps,
 * sim.cpp:751-916): border (and mid) layers take the free coefficients from
 * the vector, the layers between are fitted by the reference's transmural
 * shape interpolation (sim.cpp:91-313, nonlinearFit.h:92-168).  layer_k:
 * [n_layers][9]; violation: summed distance of free values outside
 * [k min, k max] (sim.cpp:582-596). */
int ekg_layer_aps(const ekg_ap_model *m, int n_layers, const double *genes, int n_genes,
                  double *layer_k, double *violation);

/* ---- comparison with a target ECG (host, FP64) --------------------------- */
/* `[optimization] comparison mode` (SimSettings.h:77-82). */
#define EKG_CMP_RMS 1
#define EKG_CMP_CORRELATION 2        /* 1 - Pearson r */
#define EKG_CMP_DEV_FROM_LINEAR 3
#define EKG_CMP_VECTOR_CORRELATION 4 /* 1 - a.b / (|a| |b|) */

/* Comparison value of one simulated lead against one target at offset 0,
 * over the first min(n_sim, n_target) samples (sim.cpp:640-660,
 * vectorMath.h:277-401).  target_offset is used by EKG_CMP_DEV_FROM_LINEAR
 * only.  The value is returned as computed: the caller's criterion maps a
 * negative, zero, subnormal or non-finite value to 2e10 (sim.cpp:679-685). */
int ekg_criterion(int mode, const double *sim, int64_t n_sim, const double *target, int64_t n_target,
                  double target_offset, double *value);

/* ---- comparison on the device, after the ECG reduction ------------------ */
/* The per-lead part of SimImplementation::calculateFitness (sim.cpp:600-702)
 * for a batch whose ECGs are on the device: for request r and lead
 * i < n_crit = min(n_leads, n_targets), the ekg_criterion value of the lead
 * against target i (bit for bit the host's), then with `peaks` the distance
 * |argmax(sim) - argmax(target)| of the first maxima (sim.cpp:662-670):
 * out[r][i] or out[r][2i], out[r][2i+1].  Values that are negative, zero,
 * subnormal or not finite become 2e10 (sim.cpp:679-685) unless `raw`. */
typedef struct ekg_criteria_desc {
    int mode;                      /* EKG_CMP_* */
    int peaks;                     /* 1: peak distance after each value ("peak position is criterion") */
    int raw;                       /* 1: values as computed, without the 2e10 correction */
    int n_targets;
    int64_t target_len;            /* samples per target */
    const double *targets;         /* host [n_targets][target_len], prepared as loadTargets does */
    const double *target_offsets;  /* host [n_targets]: EKG_CMP_DEV_FROM_LINEAR offsets */
} ekg_criteria_desc;
/* d_ecg: device [batch][n_leads][n_steps]; d_out: device [batch][n_crit (x2 with peaks)];
 * runs on the context's stream. */
int ekg_criteria_device(ekg_ctx *ctx, const ekg_criteria_desc *c, const double *d_ecg, int batch, int n_leads,
                        int64_t n_steps, double *d_out);
/* Simulate a batch (one launch group, as ekg_simulate_batch; every request
 * with the same lead count and steps) and return only its criteria (host
 * [batch][n_crit (x2 with peaks)]): the ECGs stay on the device.  Replaces
 * the simulate-then-compare sequence of SimImplementation::eval
 * (sim.cpp:443-491, 600-702) for batched evaluations. */
int ekg_simulate_criteria(ekg_heart *h, int batch, const ekg_sim_desc *descs, const ekg_criteria_desc *c,
                          double *criteria);

/* Target resampling of SimImplementation::loadTargets (sim.cpp:1013-1022,
 * vectorMath.h:209-227): dest[i] = linear interpolation of src at
 * i / factor + start for 1 + ceil((n - 1 - start) * factor) samples; a factor
 * of exactly 1 copies src unshifted (the reference ignores start then).
 * *n_out receives the full length; at most cap samples are written (dest
 * may be NULL with cap 0 to query the length). */
int ekg_resample(const double *src, int64_t n, double factor, int start, double *dest, int64_t cap,
                 int64_t *n_out);

/* ---- serving (continuous batching) -------------------------------------- */
/* Device-resident request loop standing in for AMS-DEMO's
 * ExternalEvaluation / PFramework job queues (ExternalEvaluation.h:95-151,
 * ParallelFramework.h:286-430): requests are admitted while their device
 * memory fits hbm_budget_bytes and a slot is free, grouped into fused
 * launches per dipole neighbourhood and contraction path (requests of several
 * hearts share a group), and overlap on the GPU with requests already
 * running. */
typedef struct ekg_server_opts {
    size_t hbm_budget_bytes;  /* device memory the server may use for requests (0 = 16 GiB) */
    int max_inflight;         /* requests on the GPU at once (0 = 1024) */
    int max_batch;            /* requests per launch group (0 = 256) */
    int n_streams;            /* concurrent launch groups (0 = 2) */
    int batch_window_us;      /* while another group runs, a group that could still grow (below
                                 max_batch and the in-flight cap, inside the budget) waits up to
                                 this long (from its oldest request) for more requests
                                 (0 = 1000 us; < 0 = launch at once) */
} ekg_server_opts;

int ekg_server_create(ekg_ctx *ctx, const ekg_server_opts *opts, ekg_server **out);
/* Stops the worker after finishing admitted requests; pending ones fail. */
void ekg_server_destroy(ekg_server *s);
/* Caches a heart (geometry + activation times) for requests; the heart must
 * stay alive while the server uses it. */
int ekg_server_add_heart(ekg_server *s, ekg_heart *h, int *heart_id);
/* Queue one simulation; returns at once with a ticket. */
int ekg_server_submit(ekg_server *s, int heart_id, const ekg_sim_desc *d, uint64_t *ticket);
/* done = 1 when the result is ready (or failed). */
int ekg_server_poll(ekg_server *s, uint64_t ticket, int *done);
/* Block until the ticket is done, copy its ECG ([n_leads][T] FP64, cap
 * doubles available) and release it.  latency_ms (optional) receives
 * submit-to-completion time. */
int ekg_server_wait(ekg_server *s, uint64_t ticket, double *ecg, size_t cap, double *latency_ms);
/* Block until the ticket is done and hand out its ECG in place: *ecg points
 * at n doubles ([n_leads][T] FP64) in the server's pinned host memory, valid
 * until ekg_server_release(ticket).  Saves the copy of ekg_server_wait for
 * callers that read the result once (a holder keeps its whole launch
 * group's buffer from reuse).  No counterpart in the reference (the
 * request loop's result hand-off, like ekg_server_wait). */
int ekg_server_wait_view(ekg_server *s, uint64_t ticket, const double **ecg, size_t *n, double *latency_ms);
int ekg_server_release(ekg_server *s, uint64_t ticket);
/* Counters: completed requests, launch groups, peak in-flight requests,
 * peak device bytes. */
int ekg_server_stats(ekg_server *s, uint64_t *completed, uint64_t *groups, int *peak_inflight,
                     size_t *peak_bytes);

#ifdef __cplusplus
}
#endif
#endif /* EKG_H_INCLUDED */
