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

// WohlfartPlus (simlib/Wohlfart.h:167-230): the nine-coefficient AP curve.
// Values come from ekg_ap_value, which keeps the reference's FP64 operation
// order; apd90 from ekg_ap_apd90.  Declared at global scope, where the
// reference's callers name it (sim.cpp:521, 540).
class WohlfartPlus {
public:
    static const size_t numParams = 9;

    WohlfartPlus() {}
    explicit WohlfartPlus(const double newK[numParams]) { setK(newK); }
    void setK(const double newK[numParams])
    {
        for (size_t i = 0; i < numParams; ++i) k_[i] = newK[i];
    }
    void setK(const std::vector<double>& newK)
    {
        for (size_t i = 0; i < newK.size() && i < numParams; ++i) k_[i] = newK[i];
    }
    const double* getK() const { return k_; }
    double* getK() { return k_; }
    // value of the curve at time t (ms after activation)
    double operator[](double t) const { return ekg_ap_value(k_, t); }
    // time at which the curve last falls through k0 + k2/10 over 0..999 ms
    // (-1 if it never does)
    double apd90() const { return ekg_ap_apd90(k_); }

private:
    double k_[numParams] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
};

// .column reader (simlib/columnFile.h:84-186, reached through the
// reference's simlib headers): equal-length columns of doubles.  Comment lines
// start with '#' or '%'; the first data line fixes the column count; reading
// stops at the first line with fewer numbers.  The guard macro is the
// reference header's, so a caller that also includes that header directly
// (main.cpp:82) gets one definition.
#ifndef COLUMNFILEL_H_INCLUDED
#define COLUMNFILEL_H_INCLUDED
typedef std::vector<double> Column;
typedef std::vector<Column> Columns;

class ColumnFile {
public:
    ColumnFile() {}
    explicit ColumnFile(const char* fname) { load(fname); }
    Columns& columns() { return cols_; }
    const Columns& columns() const { return cols_; }
    Column& operator[](size_t p) { return cols_[p]; }
    const Column& operator[](size_t p) const { return cols_[p]; }
    size_t numColumns() const { return cols_.size(); }
    size_t columnsLength() const { return cols_.empty() ? 0 : cols_[0].size(); }
    // throws std::runtime_error for other extensions, unreadable files and
    // files without a data line
    void load(const char* fname);

private:
    Columns cols_;
};
#endif

namespace SimLib {

// (z, y, x) position or displacement, the storage order of the reference's
// measuring points (simulator.cpp:377-381).
typedef std::array<double, 3> PositionVec;

// ActionPotential (simulator.h:467-480): a WohlfartPlus curve whose k8 is
// already reduced by the activation time `at`; the value at time t is
// W(t - at).
struct ActionPotential {
    WohlfartPlus wohl;
    double at = 0;

    void init(const double* wohlK, double newAt);
    void init(const ActionPotential& ap, double newAt);
    double operator()(double time) const { return wohl[time - at]; }
    double& operator[](size_t i) { return wohl.getK()[i]; }
    double operator[](size_t i) const { return wohl.getK()[i]; }
    const double* getK() const { return wohl.getK(); }
    double* getK() { return wohl.getK(); }
    size_t size() const { return WohlfartPlus::numParams; }
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
    typedef SimLib::PositionVec PositionVec;  // sim_lib.h:95
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
// saveVecElement is the reference's element type (simulator.h:184-197).
struct saveVecElement {
    std::string name, comment;
    const std::vector<double>* data = nullptr;
    size_t size() const { return data->size(); }
    double operator[](size_t n) const { return (*data)[n]; }
};
typedef saveVecElement ColumnVec;
void exportVectors(const std::vector<saveVecElement>& vec, const std::string& filename, double startTime,
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

// The reference makes the class visible at global scope too (sim_lib.h:290).
typedef SimLib::EkgSim EkgSim;
