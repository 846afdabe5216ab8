// ekgsim_app.hpp -- the evaluation behind the reference's `-sim` vector:
// simulator.ini optimisation settings (SimSettings.h), parameter vector ->
// layer APs and electrode displacements, GPU simulation, criteria and the
// output files (SimImplementation, sim.cpp:321-1042; OptimizationFunction,
// sim.cpp:1081-1127).  One Evaluator holds one heart on one GPU.
#pragma once

#include <array>
#include <memory>
#include <string>
#include <vector>

#include "sim_lib.hpp"

namespace ekgapp {

// Which files an evaluation writes (sim.h:79-85): result.column, layer
// APs, selected per-cell APs.
struct OutputSettings {
    std::vector<size_t> outputCellAps;
    bool layerAps = false;
    bool excitationSequence = false;
    bool result = false;
};

// Optimisation-side settings of simulator.ini (SimSettings.h:73-356).
struct SimSettings {
    enum CriteriaMode { mode_every_lead_is_criterium = 1, mode_leads_sum_is_criterium = 9 };
    enum ComparisonMode {
        mode_rms = 1,
        mode_correlation = 2,
        mode_normalization_offset_division_variance = 3,
        mode_vector_correlation = 4
    };

    std::vector<std::array<double, 9>> baseAps;
    std::vector<double> kMin, kMax;
    std::string interpolationTypeString;
    std::vector<int> freeKs;
    bool measuringPointsDisplacementIsInput = false;
    std::vector<int> displacementMin, displacementMax;
    std::string optimizationTargetsFname;
    ComparisonMode comparisonMode = mode_correlation;
    CriteriaMode criteriaMode = mode_every_lead_is_criterium;
    double endoEpiMinCriterionDelay = -1;
    size_t numGenerations = 100;
    size_t populationSize = 0;
    int queueSize = 1;
    double midPosition = 0.5;
    bool peakPositionIsCriterion = false;
    bool fastApproxIsCriterion = false;
    double fastApproxEpiDelay = 0;
    double fastApproxLimit = 2;

    void load(const char* fname);
};

// .column reader (ColumnFile::load, simlib/columnFile.h:125-174): leading
// '#'/'%' lines skipped, the first data line fixes the column count, rows
// are read until one does not parse.
std::vector<std::vector<double>> readColumns(const std::string& fname);

// Linear resampling (vectorMath.h:209-227): factor 1 copies the source and
// ignores startOffset.
void resample(const std::vector<double>& src, std::vector<double>& dest, double factor, int startOffset);

class Evaluator {
public:
    // Loads simulator.ini (both the simulator and optimisation parts), the
    // conduction matrix, measuring points, shape and excitation sequence,
    // and the targets, in the reference's order (sim.cpp:358-430).
    explicit Evaluator(const char* ini = "simulator.ini", int device = -1);

    OutputSettings outSettings;

    size_t numCriteria() const { return num_criteria_; }
    size_t numGenes() const { return num_wohlfart_ + num_displacement_; }
    const SimSettings& settings() const { return settings_; }
    SimLib::EkgSim& sim() { return *sim_; }

    // One evaluation (SimImplementation::eval, sim.cpp:443-491): result gets
    // the criteria, the return value is the violation.  Writes the files
    // named in outSettings.
    double eval(const std::vector<double>& solution, std::vector<double>& result);

    // Many evaluations in one batched GPU simulation, without output files;
    // results[i] and violations[i] equal eval(solutions[i]).
    void evalBatch(const std::vector<std::vector<double>>& solutions, std::vector<std::vector<double>>& results,
                   std::vector<double>& violations);

    // Gene count, criteria count and gene bounds (sim.cpp:512-578).
    void getGeneParams(size_t& numGenes, size_t& numCriteria, std::vector<double>& gMin,
                       std::vector<double>& gMax);

private:
    struct Prepared {
        std::vector<SimLib::ActionPotential> layerAps;
        std::vector<SimLib::PositionVec> points;
        double violation = 0;
        std::vector<double> approx;
        double approxCriteria = 2;
        bool simulate = true;
    };

    void loadTargets(const char* fname);
    void prepare(const std::vector<double>& solution, Prepared& p);
    double compare(const std::vector<double>& a, size_t target) const;
    // crit: per-lead values (and peak distances) computed on the device for a
    // simulated solution, in ekg_criteria_desc order (raw values); null: from `ecg`.
    void fitness(const Prepared& p, const std::vector<std::vector<double>>* ecg, std::vector<double>& result,
                 const double* crit = nullptr) const;
    void writeOutputs(const std::vector<double>& solution, const Prepared& p);

    SimSettings settings_;
    std::unique_ptr<SimLib::EkgSim> sim_;
    int interpolation_ = 0;
    size_t num_wohlfart_ = 0, num_displacement_ = 0, num_criteria_ = 0;
    std::vector<std::vector<double>> targets_;
    std::vector<double> target_offsets_;
    size_t evaluations_ = 0;
};

}  // namespace ekgapp
