/*
 * ekg_oracle.c -- FP64 CPU restatement of the ekgsim simulation core.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP product path:
 * only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may load
 * it.  The product library never links it and never falls back to it.
 *
 * Every function follows one piece of the reference (synergy-twinning/ekgsim);
 * the file:line it restates is given in its comment.  Arithmetic is IEEE FP64
 * with the same libm calls (exp, log, pow, sqrt) and, where it matters for
 * bit-exactness, the same operation order.  Build with -ffp-contract=off and
 * no -march so no fused multiply-add is formed (the reference build,
 * ekgSim.mk:53, has no -march either).
 *
 * Pinning (how each function is checked, beyond this restatement):
 *   orc_wohlfart        closed-form known answers + compiled reference simlib
 *   orc_excite_*        brute-force Bellman-Ford on small grids + compiled
 *                       reference (FP64 mode); INT mode: Bellman-Ford only
 *   orc_run_literal     compiled reference simlib (bit-exact) + invariants
 *   orc_run_regrouped   orc_run_literal (<=1e-12 rel) + invariants
 *   orc_lead_basis /    compiled reference EkgSim (bit-exact)
 *   orc_move_point
 *   orc_layer_aps_sim   compiled reference ekgSim binary (-out layer_aps)
 *   orc_pearson_crit    numpy.corrcoef
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_OK 0
#define ORC_EINVAL 1
#define ORC_ENOMEM 2
#define ORC_ENOSEED 3
#define ORC_ETRANSFER 4

/* ------------------------------------------------------------------------ */
/* Wohlfart AP                                                              */
/* ------------------------------------------------------------------------ */

/* WohlfartPlus::operator[] (simlib/Wohlfart.h:195-203), same expression tree. */
double orc_wohlfart(const double *k, double t)
{
    return (1.0 / (1.0 + exp(-k[1] * t)))
         * (k[2] * ((1.0 - k[3]) * exp(-k[4] * t) + k[3]))
         * (exp(-k[5] * t) * (1 - pow((1 + exp(-k[7] * (t - k[8])
               + log(pow(2, (k[7] / k[6])) - 1))), -(k[6] / k[7]))))
         + k[0];
}

/* ActionPotential::init + operator() (simlib/simulator.cpp:154-170):
 * the cell AP copies the layer coefficients, shifts k8 by the activation
 * time and is evaluated at (time - at). */
double orc_cell_ap(const double *layer_k, double at, double time)
{
    double k[9];
    memcpy(k, layer_k, sizeof k);
    k[8] -= at;
    return orc_wohlfart(k, time - at);
}

void orc_wohlfart_series(const double *k, const double *t, int64_t n, double *out)
{
    for (int64_t i = 0; i < n; ++i) out[i] = orc_wohlfart(k, t[i]);
}

/* ------------------------------------------------------------------------ */
/* Neighbourhoods                                                           */
/* ------------------------------------------------------------------------ */

/* Neighbourhood::create + callbacks (simlib/simulator.h:325-390).
 * active runs over {0,1,2}^3 with active[0] outermost; the stored offset is
 * dif = active - (1,1,1).  Types: 0 = 2D4 (callback4N), 1 = 2D8 (callback8N),
 * 2 = 3D4 (callback3x2N, the 8 corners), 3 = 3D8/cube (callbackCube).
 * Returns the number of offsets written to dif[*][3] (at most 26). */
int orc_neighbourhood(int type, int dif[26][3])
{
    int n = 0;
    for (int a0 = 0; a0 < 3; ++a0)
        for (int a1 = 0; a1 < 3; ++a1)
            for (int a2 = 0; a2 < 3; ++a2) {
                int d0 = abs(a0 - 1), d1 = abs(a1 - 1), d2 = abs(a2 - 1);
                int mn12 = d1 < d2 ? d1 : d2, mx12 = d1 > d2 ? d1 : d2;
                int mn = d0 < mn12 ? d0 : mn12, mx = d0 > mx12 ? d0 : mx12;
                int take = 0;
                switch (type) {
                case 0: take = (mn12 == 1 && mx12 == 1 && d0 == 0); break;
                case 1: take = (mx12 == 1 && d0 == 0); break;
                case 2: take = (mn == 1 && mx == 1); break;
                case 3: take = (mx == 1); break;
                default: return -1;
                }
                if (take) {
                    dif[n][0] = a0 - 1; dif[n][1] = a1 - 1; dif[n][2] = a2 - 1;
                    ++n;
                }
            }
    return n;
}

/* ------------------------------------------------------------------------ */
/* Excitation sequence (cellular automaton)                                 */
/* ------------------------------------------------------------------------ */

typedef struct { double time; int64_t idx; } heap_el;
typedef struct { heap_el *a; int64_t n, cap; } heap_t;

static int heap_push(heap_t *h, double time, int64_t idx)
{
    if (h->n == h->cap) {
        int64_t nc = h->cap ? 2 * h->cap : 1024;
        heap_el *na = (heap_el *)realloc(h->a, (size_t)nc * sizeof(heap_el));
        if (!na) return ORC_ENOMEM;
        h->a = na; h->cap = nc;
    }
    int64_t i = h->n++;
    while (i > 0) {
        int64_t p = (i - 1) / 2;
        if (h->a[p].time <= time) break;
        h->a[i] = h->a[p];
        i = p;
    }
    h->a[i].time = time; h->a[i].idx = idx;
    return ORC_OK;
}

static heap_el heap_pop(heap_t *h)
{
    heap_el top = h->a[0];
    heap_el last = h->a[--h->n];
    int64_t i = 0;
    for (;;) {
        int64_t c = 2 * i + 1;
        if (c >= h->n) break;
        if (c + 1 < h->n && h->a[c + 1].time < h->a[c].time) ++c;
        if (last.time <= h->a[c].time) break;
        h->a[i] = h->a[c];
        i = c;
    }
    if (h->n > 0) h->a[i] = last;
    return top;
}

/* Simulation::calculateExcitationSequence + exciteElement
 * (simlib/simulator.cpp:212-286).  Priority-queue flood from every seed at
 * time 1.0 over the 26-neighbourhood in 3D (8-neighbourhood if n0 == 1,
 * simulator.cpp:251-254).  The lag to a neighbour v of an excited voxel u is
 * transfer[layer(u)][layer(v)] * sqrt(|dif|^2) in FP64 (simulator.cpp:239-241).
 * act == 0 marks "not excited" (simulator.cpp:219); unreachable myocardium
 * and empty voxels keep 0.  transfer is row-major [rows][cols]. */
int orc_excite_fp64(int n0, int n1, int n2, const uint8_t *layer, const uint8_t *seed,
                    const double *transfer, int rows, int cols, double *act)
{
    int dif[26][3];
    int nn = orc_neighbourhood(n0 > 1 ? 3 : 1, dif);
    int64_t nvox = (int64_t)n0 * n1 * n2;
    heap_t h = {0, 0, 0};
    int64_t nseeds = 0;
    for (int64_t i = 0; i < nvox; ++i) act[i] = 0.0;
    for (int64_t i = 0; i < nvox; ++i)
        if (seed[i]) {
            if (layer[i] == 0) { free(h.a); return ORC_EINVAL; }
            if (heap_push(&h, 1.0, i)) { free(h.a); return ORC_ENOMEM; }
            ++nseeds;
        }
    if (nseeds == 0) { free(h.a); return ORC_ENOSEED; }
    while (h.n > 0) {
        heap_el e = heap_pop(&h);
        if (act[e.idx] != 0.0) continue;
        act[e.idx] = e.time;
        int64_t i2 = e.idx % n2, r = e.idx / n2;
        int64_t i1 = r % n1, i0 = r / n1;
        int lu = layer[e.idx];
        for (int k = 0; k < nn; ++k) {
            int64_t j0 = i0 - dif[k][0], j1 = i1 - dif[k][1], j2 = i2 - dif[k][2];
            if (j0 < 0 || j0 >= n0 || j1 < 0 || j1 >= n1 || j2 < 0 || j2 >= n2) continue;
            int64_t j = (j0 * n1 + j1) * n2 + j2;
            int lv = layer[j];
            if (lv == 0 || act[j] != 0.0) continue;
            if (lu >= cols || lv >= cols || lu >= rows || lv >= rows) { free(h.a); return ORC_ETRANSFER; }
            double lag = transfer[lu * cols + lv];
            int sq = dif[k][0] * dif[k][0] + dif[k][1] * dif[k][1] + dif[k][2] * dif[k][2];
            lag *= sqrt((double)sq);
            if (heap_push(&h, e.time + lag, j)) { free(h.a); return ORC_ENOMEM; }
        }
    }
    free(h.a);
    return ORC_OK;
}

/* Per-voxel integer-delay variant of the same flood (the synthetic-heart
 * configuration): seeds at 1, the lag into a neighbour v is delay[v] steps
 * for all 26 (or 8 in 2D) neighbours, 0 marks unreached.  Same queue
 * discipline as simulator.cpp:212-246 with integer keys. */
int orc_excite_int(int n0, int n1, int n2, const uint8_t *layer, const uint8_t *seed,
                   const uint8_t *delay, int32_t *act)
{
    int dif[26][3];
    int nn = orc_neighbourhood(n0 > 1 ? 3 : 1, dif);
    int64_t nvox = (int64_t)n0 * n1 * n2;
    heap_t h = {0, 0, 0};
    int64_t nseeds = 0;
    for (int64_t i = 0; i < nvox; ++i) act[i] = 0;
    for (int64_t i = 0; i < nvox; ++i)
        if (seed[i]) {
            if (layer[i] == 0) { free(h.a); return ORC_EINVAL; }
            if (heap_push(&h, 1.0, i)) { free(h.a); return ORC_ENOMEM; }
            ++nseeds;
        }
    if (nseeds == 0) { free(h.a); return ORC_ENOSEED; }
    while (h.n > 0) {
        heap_el e = heap_pop(&h);
        if (act[e.idx] != 0) continue;
        act[e.idx] = (int32_t)e.time;
        int64_t i2 = e.idx % n2, r = e.idx / n2;
        int64_t i1 = r % n1, i0 = r / n1;
        for (int k = 0; k < nn; ++k) {
            int64_t j0 = i0 - dif[k][0], j1 = i1 - dif[k][1], j2 = i2 - dif[k][2];
            if (j0 < 0 || j0 >= n0 || j1 < 0 || j1 >= n1 || j2 < 0 || j2 >= n2) continue;
            int64_t j = (j0 * n1 + j1) * n2 + j2;
            if (layer[j] == 0 || act[j] != 0) continue;
            if (heap_push(&h, e.time + (double)delay[j], j)) { free(h.a); return ORC_ENOMEM; }
        }
    }
    free(h.a);
    return ORC_OK;
}

/* Brute-force Bellman-Ford fixpoint of the FP64 flood: sweep until no value
 * changes, a(v) = min over excited neighbours u of fl(a(u) + w(u,v)).  Used
 * only to pin orc_excite_fp64 on small grids (an independent algorithm that
 * must land on the same bits). */
int orc_excite_fp64_bf(int n0, int n1, int n2, const uint8_t *layer, const uint8_t *seed,
                       const double *transfer, int rows, int cols, double *act)
{
    int dif[26][3];
    int nn = orc_neighbourhood(n0 > 1 ? 3 : 1, dif);
    int64_t nvox = (int64_t)n0 * n1 * n2;
    int64_t nseeds = 0;
    for (int64_t i = 0; i < nvox; ++i) {
        act[i] = seed[i] ? 1.0 : INFINITY;
        if (seed[i]) { if (!layer[i]) return ORC_EINVAL; ++nseeds; }
    }
    if (!nseeds) return ORC_ENOSEED;
    for (int changed = 1; changed;) {
        changed = 0;
        for (int64_t v = 0; v < nvox; ++v) {
            if (!layer[v]) continue;
            int64_t i2 = v % n2, r = v / n2, i1 = r % n1, i0 = r / n1;
            double best = act[v];
            for (int k = 0; k < nn; ++k) {
                int64_t j0 = i0 + dif[k][0], j1 = i1 + dif[k][1], j2 = i2 + dif[k][2];
                if (j0 < 0 || j0 >= n0 || j1 < 0 || j1 >= n1 || j2 < 0 || j2 >= n2) continue;
                int64_t u = (j0 * n1 + j1) * n2 + j2;
                if (!layer[u] || act[u] == INFINITY) continue;
                if (layer[u] >= rows || layer[u] >= cols || layer[v] >= rows || layer[v] >= cols)
                    return ORC_ETRANSFER;
                int sq = dif[k][0] * dif[k][0] + dif[k][1] * dif[k][1] + dif[k][2] * dif[k][2];
                double lag = transfer[layer[u] * cols + layer[v]];
                lag *= sqrt((double)sq);
                double cand = act[u] + lag;
                if (cand < best) best = cand;
            }
            if (best < act[v]) { act[v] = best; changed = 1; }
        }
    }
    for (int64_t i = 0; i < nvox; ++i)
        if (act[i] == INFINITY) act[i] = 0.0;
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* Measuring points                                                         */
/* ------------------------------------------------------------------------ */

/* normalize() (copyOfLibs/Pattern/Hypermatrix.h:264-270): multiply by 1/l. */
static void orc_normalize(double v[3])
{
    double l = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    double s = (l != 0) ? 1.0 / l : 0.0;
    v[0] *= s; v[1] *= s; v[2] *= s;
}

/* EkgSim::loadMeasuringPoints U/V basis (simlib/sim_lib.h:167-193).
 * p is in the reference's (z,y,x) storage order; x-hat is (0,0,1) there.
 * U = normalize(-(x-hat cross p)), V = normalize((x-hat cross p) cross p). */
void orc_lead_basis(const double *p, double *U, double *V)
{
    const double x[3] = {0, 0, 1};
    double u[3], v[3];
    u[0] = x[1] * p[2] - x[2] * p[1];
    u[1] = x[2] * p[0] - x[0] * p[2];
    u[2] = x[0] * p[1] - x[1] * p[0];
    v[0] = u[1] * p[2] - u[2] * p[1];
    v[1] = u[2] * p[0] - u[0] * p[2];
    v[2] = u[0] * p[1] - u[1] * p[0];
    u[0] *= -1; u[1] *= -1; u[2] *= -1;
    orc_normalize(u);
    orc_normalize(v);
    memcpy(U, u, sizeof u);
    memcpy(V, v, sizeof v);
}

/* EkgSim::moveMeasuringPoints (simlib/sim_lib.h:197-208):
 * out = (p + du*U) + dv*V, element-wise. */
void orc_move_point(const double *p, const double *U, const double *V,
                    double du, double dv, double *out)
{
    for (int k = 0; k < 3; ++k) out[k] = (p[k] + U[k] * du) + V[k] * dv;
}

/* ------------------------------------------------------------------------ */
/* ECG simulation                                                           */
/* ------------------------------------------------------------------------ */

static int64_t orc_total_steps(double length, double dt)
{
    return (int64_t)ceil(length / dt);   /* simulator.cpp:470 */
}

int64_t orc_num_steps(double length, double dt) { return orc_total_steps(length, dt); }

/* Simulation::run, literal loop (simlib/simulator.cpp:452-550).
 *   layer[n0*n1*n2]     0 = empty, l >= 1 uses layer_k[(l-1)*9 .. +9]
 *   act[]               activation time per voxel (0 for unreached)
 *   pos[n_leads*3]      measuring points in (z,y,x) order
 *   ecg[n_leads*T]      output, T = ceil(length/dt)
 * The voxel at model index u sits at bordered index u+1 (simulator.cpp:458-463),
 * raster order with the last index fastest; per voxel: cell AP series, dipole
 * over the neighbourhood, projection m = d/|d|^3 onto each lead. */
int orc_run_literal(int n0, int n1, int n2, const uint8_t *layer, const double *act,
                    int n_layers, const double *layer_k, int nhood_type,
                    double start, double length, double dt,
                    int n_leads, const double *pos, double *ecg)
{
    int dif[26][3];
    int nn = orc_neighbourhood(nhood_type, dif);
    if (nn < 0) return ORC_EINVAL;
    int64_t T = orc_total_steps(length, dt);
    int64_t nvox = (int64_t)n0 * n1 * n2;
    for (int64_t i = 0; i < nvox; ++i)
        if (layer[i] > n_layers) return ORC_EINVAL;
    int b0 = n0 + 2, b1 = n1 + 2, b2 = n2 + 2;
    int64_t nb = (int64_t)b0 * b1 * b2;
    uint8_t *bl = (uint8_t *)calloc((size_t)nb, 1);
    double *ba = (double *)calloc((size_t)nb, sizeof(double));
    double *cell = (double *)malloc((size_t)T * sizeof(double));
    double *dip = (double *)malloc((size_t)T * 3 * sizeof(double));
    if (!bl || !ba || !cell || !dip) { free(bl); free(ba); free(cell); free(dip); return ORC_ENOMEM; }
    for (int i0 = 0; i0 < n0; ++i0)
        for (int i1 = 0; i1 < n1; ++i1)
            for (int i2 = 0; i2 < n2; ++i2) {
                int64_t s = ((int64_t)i0 * n1 + i1) * n2 + i2;
                int64_t d = ((int64_t)(i0 + 1) * b1 + (i1 + 1)) * b2 + (i2 + 1);
                bl[d] = layer[s];
                ba[d] = act[s];
            }
    for (int64_t i = 0; i < (int64_t)n_leads * T; ++i) ecg[i] = 0.0;
    double total_time = length + start;   /* simulator.cpp:481 */
    int rc = ORC_OK;
    for (int i0 = 1; i0 <= n0 && rc == ORC_OK; ++i0)
        for (int i1 = 1; i1 <= n1 && rc == ORC_OK; ++i1)
            for (int i2 = 1; i2 <= n2; ++i2) {
                int64_t b = ((int64_t)i0 * b1 + i1) * b2 + i2;
                if (bl[b] == 0) continue;
                const double *ku = layer_k + (bl[b] - 1) * 9;
                double sim_time = start;
                for (int64_t j = 0; j < T; ++j) {
                    cell[j] = orc_cell_ap(ku, ba[b], sim_time);
                    sim_time += dt;
                }
                for (int64_t j = 0; j < 3 * T; ++j) dip[j] = 0.0;
                for (int k = 0; k < nn; ++k) {
                    int64_t b2i = ((int64_t)(i0 - dif[k][0]) * b1 + (i1 - dif[k][1])) * b2 + (i2 - dif[k][2]);
                    if (bl[b2i] == 0) continue;
                    const double *kv = layer_k + (bl[b2i] - 1) * 9;
                    int64_t next = 0;
                    for (double st = start; st < total_time; st += dt) {
                        if (next >= T) { rc = ORC_EINVAL; break; }
                        double vd = orc_cell_ap(kv, ba[b2i], st) - cell[next];
                        dip[3 * next + 0] += (double)dif[k][0] * vd;
                        dip[3 * next + 1] += (double)dif[k][1] * vd;
                        dip[3 * next + 2] += (double)dif[k][2] * vd;
                        ++next;
                    }
                    if (rc) break;
                }
                if (rc) break;
                for (int i = 0; i < n_leads; ++i) {
                    double m[3];
                    m[0] = pos[3 * i + 0] - (double)i0;
                    m[1] = pos[3 * i + 1] - (double)i1;
                    m[2] = pos[3 * i + 2] - (double)i2;
                    double l = m[0] * m[0];
                    l += m[1] * m[1];
                    l += m[2] * m[2];
                    double s = 1.0 / (l * sqrt(l));
                    m[0] *= s; m[1] *= s; m[2] *= s;
                    double *e = ecg + (int64_t)i * T;
                    for (int64_t t = 0; t < T; ++t) {
                        double r = 0;
                        r += m[0] * dip[3 * t + 0];
                        r += m[1] * dip[3 * t + 1];
                        r += m[2] * dip[3 * t + 2];
                        e[t] += r;
                    }
                }
            }
    free(bl); free(ba); free(cell); free(dip);
    return rc;
}

/* Lead weight of the voxel at model index (i0,i1,i2) for a lead at p,
 * m = (p - (u+1)) * (1/(l*sqrt(l))) -- simulator.cpp:530-533 with
 * sqrLength/pow3Length of Hypermatrix.h:250-261. */
static void orc_m(const double *p, int64_t i0, int64_t i1, int64_t i2, double m[3])
{
    m[0] = p[0] - (double)(i0 + 1);
    m[1] = p[1] - (double)(i1 + 1);
    m[2] = p[2] - (double)(i2 + 1);
    double l = m[0] * m[0];
    l += m[1] * m[1];
    l += m[2] * m[2];
    double s = 1.0 / (l * sqrt(l));
    m[0] *= s; m[1] *= s; m[2] *= s;
}

/* Regrouped lead field.  Summing simulator.cpp:509-537 over all voxels and
 * collecting the coefficient of each cell potential V_x gives
 *     ECG_i(t) = sum_x G_i(x) V_x(t),
 *     G_i(x)   = sum_{d in N, x+d myocardial} (m_i(x+d) + m_i(x)) . d
 * because every neighbourhood N is symmetric (d in N <=> -d in N).
 * Output G[v*n_leads + i] for the myocardial voxels in raster order;
 * returns the number of myocardial voxels (or -1 on error). */
int64_t orc_lead_field(int n0, int n1, int n2, const uint8_t *layer, int nhood_type,
                       int n_leads, const double *pos, double *G)
{
    int dif[26][3];
    int nn = orc_neighbourhood(nhood_type, dif);
    if (nn < 0) return -1;
    int64_t v = 0;
    for (int64_t i0 = 0; i0 < n0; ++i0)
        for (int64_t i1 = 0; i1 < n1; ++i1)
            for (int64_t i2 = 0; i2 < n2; ++i2) {
                int64_t x = (i0 * n1 + i1) * n2 + i2;
                if (!layer[x]) continue;
                for (int i = 0; i < n_leads; ++i) {
                    double mx[3], my[3], g = 0.0;
                    orc_m(pos + 3 * i, i0, i1, i2, mx);
                    for (int k = 0; k < nn; ++k) {
                        int64_t j0 = i0 + dif[k][0], j1 = i1 + dif[k][1], j2 = i2 + dif[k][2];
                        if (j0 < 0 || j0 >= n0 || j1 < 0 || j1 >= n1 || j2 < 0 || j2 >= n2) continue;
                        if (!layer[(j0 * n1 + j1) * n2 + j2]) continue;
                        orc_m(pos + 3 * i, j0, j1, j2, my);
                        g += (my[0] + mx[0]) * dif[k][0] + (my[1] + mx[1]) * dif[k][1]
                           + (my[2] + mx[2]) * dif[k][2];
                    }
                    G[v * n_leads + i] = g;
                }
                ++v;
            }
    return v;
}

/* Fast FP64 oracle: ECG_i(t_j) = sum_x G_i(x) V_x(t_j), t_j = start + j*dt by
 * repeated addition (simulator.cpp:496-500).  Voxels are split into 256 fixed
 * contiguous chunks whose partial sums are added in chunk order, so the result
 * does not depend on the thread count. */
int orc_run_regrouped(int n0, int n1, int n2, const uint8_t *layer, const double *act,
                      int n_layers, const double *layer_k, int nhood_type,
                      double start, double length, double dt,
                      int n_leads, const double *pos, double *ecg)
{
    int64_t nvox = (int64_t)n0 * n1 * n2;
    int64_t V = 0;
    for (int64_t i = 0; i < nvox; ++i) {
        if (layer[i] > n_layers) return ORC_EINVAL;
        if (layer[i]) ++V;
    }
    int64_t T = orc_total_steps(length, dt);
    double *G = (double *)malloc((size_t)(V ? V : 1) * n_leads * sizeof(double));
    int64_t *vx = (int64_t *)malloc((size_t)(V ? V : 1) * sizeof(int64_t));
    double *tj = (double *)malloc((size_t)(T ? T : 1) * sizeof(double));
    enum { NCH = 256 };
    double *part = (double *)calloc((size_t)NCH * n_leads * (T ? T : 1), sizeof(double));
    if (!G || !vx || !tj || !part) { free(G); free(vx); free(tj); free(part); return ORC_ENOMEM; }
    if (orc_lead_field(n0, n1, n2, layer, nhood_type, n_leads, pos, G) < 0) {
        free(G); free(vx); free(tj); free(part); return ORC_EINVAL;
    }
    int64_t w = 0;
    for (int64_t i = 0; i < nvox; ++i) if (layer[i]) vx[w++] = i;
    double st = start;
    for (int64_t j = 0; j < T; ++j) { tj[j] = st; st += dt; }
    #pragma omp parallel for schedule(dynamic, 1)
    for (int c = 0; c < NCH; ++c) {
        int64_t lo = V * c / NCH, hi = V * (c + 1) / NCH;
        double *pc = part + (int64_t)c * n_leads * T;
        double *vt = (double *)malloc((size_t)(T ? T : 1) * sizeof(double));
        for (int64_t u = lo; u < hi; ++u) {
            int64_t x = vx[u];
            const double *ku = layer_k + (layer[x] - 1) * 9;
            for (int64_t j = 0; j < T; ++j) vt[j] = orc_cell_ap(ku, act[x], tj[j]);
            for (int i = 0; i < n_leads; ++i) {
                double g = G[u * n_leads + i];
                double *p = pc + (int64_t)i * T;
                for (int64_t j = 0; j < T; ++j) p[j] += g * vt[j];
            }
        }
        free(vt);
    }
    for (int64_t i = 0; i < (int64_t)n_leads * T; ++i) {
        double s = 0.0;
        for (int c = 0; c < NCH; ++c) s += part[(int64_t)c * n_leads * T + i];
        ecg[i] = s;
    }
    free(G); free(vx); free(tj); free(part);
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* Parameter vector -> layer APs                                            */
/* ------------------------------------------------------------------------ */

/* getViolation (sim.cpp:582-596): distance of k[i] outside [kmin, kmax]. */
static double orc_violation_one(const double *k, const double *kmin, const double *kmax, int i)
{
    if (k[i] < kmin[i]) return kmin[i] - k[i];
    if (k[i] > kmax[i]) return k[i] - kmax[i];
    return 0;
}

/* SimImplementation::simUsingBorderAndMidAps for a 3-layer heart
 * (sim.cpp:825-880).  free_k must be sorted ascending (std::set<char>
 * order).  Layer 1 starts from base_first, layers mid and last from
 * base_last; with 3 layers and mid position p, the mid index is
 * floor(p*2 + 0.5) which is 1 for p in [0.25, 0.75).  For n_layers == 3 and
 * mid index 1 no interpolation runs (sim.cpp:884-902).  Writes
 * layer_k[3*9] and the summed violation. */
int orc_layer_aps_sim3(const double *base_first, const double *base_last,
                       const int *free_k, int n_free, double mid_position,
                       const double *kmin, const double *kmax,
                       const double *x, double *layer_k, double *violation)
{
    size_t mid = (size_t)floor(mid_position * (3 - 1) + 0.5);
    if (mid != 1) return ORC_EINVAL;
    int si = 0;
    double viol = 0.0;
    const double *bases[3] = {base_first, base_last, base_last};
    int slot[3] = {0, (int)mid, 2};
    for (int a = 0; a < 3; ++a) {
        double k[9];
        memcpy(k, bases[a], sizeof k);
        for (int i = 0; i < 9; ++i) {
            int is_free = 0;
            for (int f = 0; f < n_free; ++f) if (free_k[f] == i) is_free = 1;
            if (is_free) {
                k[i] = x[si++];
                viol += orc_violation_one(k, kmin, kmax, i);
            }
        }
        memcpy(layer_k + 9 * slot[a], k, sizeof k);
    }
    *violation = viol;
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* Criteria                                                                 */
/* ------------------------------------------------------------------------ */

/* statisticalCorrelationCoeff with offset 0 (vectorMath.h:286-317) and the
 * "1 - r" of comparison mode 2 (sim.cpp:650-652), including the 2e10
 * replacement of a negative or non-normal criterion (sim.cpp:674-681). */
double orc_pearson_crit(const double *a, int64_t na, const double *b, int64_t nb)
{
    int64_t len = na < nb ? na : nb;
    double inv = 1.0 / (double)len;
    double ma = 0, mb = 0, va = 0, vb = 0, cv = 0;
    for (int64_t i = 0; i < len; ++i) ma += a[i];
    ma = ma * inv;
    for (int64_t i = 0; i < len; ++i) va += (a[i] - ma) * (a[i] - ma);
    va *= inv;
    for (int64_t i = 0; i < len; ++i) mb += b[i];
    mb = mb * inv;
    for (int64_t i = 0; i < len; ++i) vb += (b[i] - mb) * (b[i] - mb);
    vb *= inv;
    for (int64_t i = 0; i < len; ++i) cv += (a[i] - ma) * (b[i] - mb);
    cv = cv / (double)len;
    double r = cv / (sqrt(va) * sqrt(vb));
    double c = 1.0 - r;
    if (c < 0 || !isfinite(c) || !isnormal(c)) c = 2e10;
    return c;
}
