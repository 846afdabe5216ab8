// ref_harness.cpp -- C entry points around the reference simlib, compiled
// from the reference's own sources (oracle/Makefile, target `ref`).  Test
// infrastructure only: it validates the FP64 restatement in ekg_oracle.c and
// produces golden vectors.  Nothing in the product links or loads it.
//
// Each function drives the reference classes through their public API in the
// order the reference application uses them (sim.cpp:362-431, sim_lib.h).
#include "sim_lib.h"

#include <cstdint>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

using namespace SimLib;

namespace {

std::string g_err;

void build_shape(ShapeMatrix& s, int n0, int n1, int n2, const uint8_t* layer, const uint8_t* seed)
{
    MatrixIndex sz;
    sz[0] = n0; sz[1] = n1; sz[2] = n2;
    s.resize(sz);
    for (int64_t i = 0; i < (int64_t)n0 * n1 * n2; ++i) {
        size_t l = layer[i];
        if (seed && seed[i]) l += ShapeElement::layerStartingPoint;
        s.elementAtIndex(i) = ShapeElement(l);
    }
}

void write_transfer(const char* path, const double* t, int rows, int cols)
{
    std::ofstream f(path);
    f << "Conduction matrix\n2d " << cols << " x " << rows << " tab\n";
    f.precision(17);
    for (int r = 0; r < rows; ++r) {
        for (int c = 0; c < cols; ++c) {
            if (c) f << '\t';
            f << t[r * cols + c];
        }
        f << '\n';
    }
}

void set_nhood(Simulation& sim, int type)
{
    switch (type) {
    case 0: sim.neighbourhood().create(&Neighbourhood::callback4N); break;
    case 1: sim.neighbourhood().create(&Neighbourhood::callback8N); break;
    case 2: sim.neighbourhood().create(&Neighbourhood::callback3x2N); break;
    default: sim.neighbourhood().create(&Neighbourhood::callbackCube); break;
    }
}

}  // namespace

extern "C" {

const char* ref_last_error() { return g_err.c_str(); }

double ref_wohlfart(const double* k, double t)
{
    WohlfartPlus w(k);
    return w[t];
}

double ref_cell_ap(const double* layer_k, double at, double time)
{
    ActionPotential ap;
    ap.init(layer_k, at);
    return ap(time);
}

// Transfer matrix is written to tmp_path in .matrix format and loaded through
// Simulation::loadTransferMatrix before loadShape, as sim.cpp:370-376 does.
int ref_excite(int n0, int n1, int n2, const uint8_t* layer, const uint8_t* seed,
               const double* transfer, int rows, int cols, const char* tmp_path, double* act)
{
    try {
        write_transfer(tmp_path, transfer, rows, cols);
        Simulation sim;
        sim.loadTransferMatrix(tmp_path);
        ShapeMatrix s;
        build_shape(s, n0, n1, n2, layer, seed);
        sim.loadShape(s);
        sim.calculateExcitationSequence("");
        const ShapeMatrix& out = sim.getShape();
        for (int64_t i = 0; i < (int64_t)n0 * n1 * n2; ++i) act[i] = out.elementAtIndex(i).excitationDelay;
        return 0;
    } catch (std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

// Simulation::run on a heart whose activation times are given (the layer
// count must equal the largest layer, simulator.cpp:562).
int ref_run(int n0, int n1, int n2, const uint8_t* layer, const double* act, int n_layers,
            const double* layer_k, int nhood, double start, double length, double dt,
            int n_leads, const double* pos, double* ecg, int64_t ecg_cap)
{
    try {
        Simulation sim;
        ShapeMatrix s;
        build_shape(s, n0, n1, n2, layer, nullptr);
        sim.loadShape(s);
        for (int64_t i = 0; i < (int64_t)n0 * n1 * n2; ++i)
            sim.getShape().elementAtIndex(i).excitationDelay = act[i];
        set_nhood(sim, nhood);
        sim.setup(dt);
        std::vector<Simulation::PositionVec> pts(n_leads);
        for (int i = 0; i < n_leads; ++i)
            for (int k = 0; k < 3; ++k) pts[i][k] = pos[3 * i + k];
        sim.setMeasuringPoints(pts);
        std::vector<ActionPotential> aps(n_layers);
        for (int l = 0; l < n_layers; ++l) aps[l].init(layer_k + 9 * l, 0);
        sim.setApsDestructive(aps);
        sim.run(start, length);
        int64_t w = 0;
        for (int i = 0; i < n_leads; ++i) {
            const std::vector<double>& m = sim.getMeasurment(i).getMeasurement();
            for (size_t t = 0; t < m.size(); ++t) {
                if (w >= ecg_cap) { g_err = "ecg buffer too small"; return 2; }
                ecg[w++] = m[t];
            }
        }
        return 0;
    } catch (std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

// Full facade pipeline driven by an ini file (paths inside it are relative to
// the current directory): EkgSim loadSettings / loadTransferMatrix /
// loadMeasuringPoints / loadShape / simExcitationSequence / applySettings /
// moveMeasuringPoints / setApsDestructive / run, as sim.cpp:362-491 orders
// them.  disp holds (du, dv) per lead; layer_k holds one 9-vector per layer.
int ref_ekgsim_pipeline(const char* ini, const double* disp, int n_leads, const double* layer_k,
                        int n_layers, double* ecg, int64_t ecg_cap, double* act, int64_t act_cap)
{
    try {
        EkgSim sim;
        sim.loadSettings(ini);
        sim.loadTransferMatrix();
        sim.loadMeasuringPoints();
        sim.loadShape();
        sim.simExcitationSequence();
        sim.applySettings();
        if (disp) {
            std::vector<EkgSim::PositionVec> d(n_leads);
            for (int i = 0; i < n_leads; ++i) {
                d[i][0] = disp[2 * i];
                d[i][1] = disp[2 * i + 1];
                d[i][2] = 0;
            }
            sim.moveMeasuringPoints(d);
        }
        std::vector<ActionPotential> aps(n_layers);
        for (int l = 0; l < n_layers; ++l) aps[l].init(layer_k + 9 * l, 0);
        sim.setApsDestructive(aps);
        if (act) {
            const ShapeMatrix& s = sim.getModelShape();
            int64_t n = (int64_t)s.size()[0] * s.size()[1] * s.size()[2];
            if (n > act_cap) { g_err = "act buffer too small"; return 2; }
            for (int64_t i = 0; i < n; ++i) act[i] = s.elementAtIndex(i).excitationDelay;
        }
        sim.run();
        int64_t w = 0;
        for (size_t i = 0; i < sim.numMeasurements(); ++i) {
            const std::vector<double>& m = sim.getMeasurement(i);
            for (size_t t = 0; t < m.size(); ++t) {
                if (w >= ecg_cap) { g_err = "ecg buffer too small"; return 2; }
                ecg[w++] = m[t];
            }
        }
        return 0;
    } catch (std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

}  // extern "C"
