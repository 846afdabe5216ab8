// sim_lib.hpp -- C++ facade with the interface of the reference's
// SimLib::EkgSim (simlib/sim_lib.h:93-287), running the simulation on the GPU
// through the C ABI of ekg.h.  A caller written against the reference's class
// (sim.cpp's SimImplementation, a custom driver) keeps its call sequence:
//
//   SimLib::EkgSim s;
//   s.loadSettings("simulator.ini");  s.applySettings();
//   s.loadTransferMatrix();  s.loadMeasuringPoints();  s.loadShape();
//   s.simExcitationSequence();            // GPU flood or a .matrix file
//   s.setApsDestructive(layerAps);        // one AP per layer
//   s.run();                              // GPU ECG
//   s.getMeasurement(i);  s.saveMeasurement("comment");
//
// Differences from the reference class, all in what it stores, not in what it
// computes: the shape is kept as flat (z,y,x) arrays (ModelShape) instead of
// a Hypermatrix of ShapeElement; getAps() builds the per-cell AP classes of
// Simulation::setApIndices (simulator.cpp:561-621) only when asked for.
// Errors are std::runtime_error, as in the reference (simulator.cpp:204, 277;
// matrix.h:141-243).
#pragma once

#include <array>
#include <cstdint>
#include <ostream>
#include <string>
#include <vector>

#include "ekg.h"

namespace ekgapp {
class IniFile;
}

namespace SimLib {

// (z, y, x) position or displacement, the storage order of the reference's
// measuring points (simulator.cpp:377-381).
typedef std::array<double, 3> PositionVec;

// ActionPotential (simulator.h:467-480): WohlfartPlus coefficients with
// k8 already reduced by the activation time `at`; the value at time t is
// W(t - at).
struct ActionPotential {
    double k[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double at = 0;

    void init(const double* wohlK, double newAt);
    void init(const ActionPotential& ap, double newAt);
    double operator()(double time) const { return ekg_ap_value(k, time - at); }
    double& operator[](size_t i) { return k[i]; }
    double operator[](size_t i) const { return k[i]; }
    const double* getK() const { return k; }
    double* getK() { return k; }
    size_t size() const { return 9; }
};

// Simulator settings (simulator.h:60-146), read from the [model],
// [action potentials], [output files] and [simulation] sections.
struct Settings {
    std::string inputShapeFilename;
    std::string inputExcitationSequenceFilename;
    std::string inputPointsFilename;
    std::string inputTransferFilename;
    std::string inputActionPotentials;
    std::string inputActionPotentialsFilter;
    double inputActionPotentialsTimeStep = 1;
    double inputActionPotentialsFilterParam = 0;
    std::string inputActionPotentialsFunction;
    double inputActionPotentialsScale = 0;
    std::string outputExcitationSequence;
    std::string outputApFilename;
    std::string outputFilename;
    double simulationLength = 800;
    double simulationStart = 0;
    std::string neighbourhoodType;
    double simulationTimeStep = 10;

    void loadFromIni(const ekgapp::IniFile& ini);
};

// Heart model on the host: layers (0 empty), excitation seeds and the
// activation time of every voxel, (z,y,x) with x fastest.
struct ModelShape {
    int size[3] = {0, 0, 0};  // {z, y, x}
    std::vector<uint8_t> layer;
    std::vector<uint8_t> seed;
    std::vector<double> activation;
    size_t voxels() const { return (size_t)size[0] * size[1] * size[2]; }
};

class EkgSim {
public:
    // device < 0: $EKG_DEVICE, else GPU 0.
    explicit EkgSim(int device = -1);
    ~EkgSim();

    void loadSettings(const char* fname = "settings.ini");
    Settings& getSettings() { return settings_; }
    const Settings& getSettings() const { return settings_; }
    void applySettings();

    void setApsDestructive(std::vector<ActionPotential>& destructible);
    const std::vector<ActionPotential>& getAps() const;

    void loadTransferMatrix();
    void loadMeasuringPoints();
    void moveMeasuringPoints(const std::vector<PositionVec>& uvwDisplacement);
    void loadShape();
    size_t requiredAps() const { return target_aps_; }
    void simExcitationSequence();
    void printSettings();
    void printMessages(std::ostream&) {}

    void run();
    void runApproximation(double delay, std::vector<double>& result);
    const std::vector<double>& getMeasurement(size_t n) const;
    size_t numMeasurements() const { return points_.size(); }
    void saveMeasurement(const std::string& comment = "");
    void saveMeasurementAs(const char* fname, const std::string& comment = "");

    const ModelShape& getModelShape() const { return shape_; }

    // ---- beyond the reference interface ----
    // Measuring points as currently placed ((z,y,x), displaced).
    const std::vector<PositionVec>& measuringPoints() const { return points_; }
    const std::vector<PositionVec>& originalMeasuringPoints() const { return original_; }
    // Dipole neighbourhood selected by applySettings (EKG_NHOOD_*).
    int neighbourhood() const { return nhood_; }
    ekg_ctx* context() const { return ctx_; }
    ekg_heart* heart() const { return heart_; }
    // Layer coefficients the next run() uses, [requiredAps()][9].
    std::vector<double> layerCoefficients() const;

private:
    EkgSim(const EkgSim&) = delete;
    void operator=(const EkgSim&) = delete;

    void check(int rc, const char* what) const;
    void saveTo(const std::string& fname, const std::string& comment);
    void buildCellAps() const;

    Settings settings_;
    ekg_ctx* ctx_ = nullptr;
    ekg_heart* heart_ = nullptr;
    int nhood_ = -1;
    double time_step_ = 1.0;
    double start_time_ = 0.0;
    size_t target_aps_ = 12;  // the reference's default before loadShape (simulator.h:532)
    std::vector<double> transfer_;
    int transfer_rows_ = 0, transfer_cols_ = 0;
    ModelShape shape_;
    bool excited_ = false;
    std::vector<PositionVec> original_, points_, u_, v_;
    std::vector<std::vector<double>> measurements_;
    // APs as the reference's Simulation holds them: one per layer until
    // run(), the per-cell classes after it (built lazily).
    mutable std::vector<ActionPotential> aps_;
    std::vector<ActionPotential> layer_aps_;
    mutable bool aps_are_cells_ = false, cells_built_ = false;
};

// .column writer of the reference (simulator.h:186-236): "#Comment: <c>",
// one "# name : comment" line per commented vector, "# ", a header line and
// one row per sample, time "%10.5f" when timeStep > 0, values in "%10e".
struct ColumnVec {
    std::string name, comment;
    const std::vector<double>* data = nullptr;
};
void exportVectors(const std::vector<ColumnVec>& vec, const std::string& filename, double startTime,
                   double timeStep, const std::string& comment = "");

// .matrix readers/writers (matrix.h:124-248, simulator.cpp:71-106).
void importShape(const std::string& filename, ModelShape& shape);
void importTransfer(const std::string& filename, std::vector<double>& t, int& rows, int& cols);
void importExcitation(const std::string& filename, ModelShape& shape);
void exportShape(const ModelShape& shape, const std::string& filename, bool excitation);
// The built-in 2D ring heart used when no shape file is given
// (simulator.h:600-644).
void generateTestShape(ModelShape& shape);

}  // namespace SimLib
