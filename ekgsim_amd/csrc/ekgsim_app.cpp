// ekgsim_app.cpp -- see ekgsim_app.hpp.
#include "ekgsim_app.hpp"

#include <algorithm>
#include <cmath>
#include <fstream>
#include <iostream>
#include <set>
#include <sstream>
#include <stdexcept>

#include "ini.hpp"

namespace ekgapp {

namespace {

template <class T>
std::string join(const std::vector<T>& v, const char* open, const char* close)
{
    std::ostringstream s;
    s << open;
    for (size_t i = 0; i < v.size(); ++i) s << (i ? "," : "") << v[i];
    s << close;
    return s.str();
}

// "<a,b,c>": the vector print of the reference's tools (VectorArithmetics.h:75-83).
template <class T>
std::string angle(const std::vector<T>& v)
{
    return join(v, "<", ">");
}

}  // namespace

// ---- settings ----------------------------------------------------------------

void SimSettings::load(const char* fname)
{
    std::cerr << "***** loading additional simulation parameters *************\n";
    IniFile ini(fname);
    if (ini.found())
        std::cerr << " from file: " << fname << "\n";
    else
        std::cerr << " file " << fname << "  not found, \n";

    const int ap = ini.section("wohlfart ap");
    for (size_t i = 0;; ++i) {
        std::vector<double> k;
        if (ini.get_array(k, "base ap " + std::to_string(i + 1), ap, 9) && k.size() == 9) {
            std::array<double, 9> a;
            std::copy(k.begin(), k.end(), a.begin());
            baseAps.push_back(a);
        } else {
            break;
        }
    }
    if (baseAps.size() < 2) throw std::runtime_error("need at least 2 base APs");
    if (!ini.get_array(kMin, "k min", ap, 9) || kMin.size() != 9) throw std::runtime_error("could not read kMin");
    if (!ini.get_array(kMax, "k max", ap, 9) || kMax.size() != 9) throw std::runtime_error("could not read kMax");
    ini.get(interpolationTypeString, "interpolation", ap);
    midPosition = 0.5;
    ini.get(midPosition, "mid AP position", ap);
    if (midPosition > 1.0 || midPosition < 0.0) {
        std::cerr << " warning, mid AP position must be in range [0..1] but is set to " << midPosition
                  << " in settings; using default value of 0.5 instead\n";
        midPosition = 0.5;
    }
    ini.get_array(freeKs, "free k", ap);

    const int mp = ini.section("measuring points");
    if (!ini.get_array(displacementMin, "displacement min", mp))
        throw std::runtime_error("could not read displacement min");
    if (!ini.get_array(displacementMax, "displacement max", mp))
        throw std::runtime_error("could not read displacement max");

    const int opt = ini.section("optimization");
    ini.get(optimizationTargetsFname, "targets filename", opt);
    int crit = 1;
    ini.get(crit, "mode", opt);
    if (crit != mode_every_lead_is_criterium && crit != mode_leads_sum_is_criterium)
        throw std::runtime_error("invalid optimization mode");
    criteriaMode = (CriteriaMode)crit;
    endoEpiMinCriterionDelay = -1;
    ini.get(endoEpiMinCriterionDelay, "endo-epi minimization criterion epi delay", opt);
    int flag = 0;
    ini.get(flag, "peak position is criterion", opt);
    peakPositionIsCriterion = flag > 0;
    measuringPointsDisplacementIsInput = false;
    ini.get(measuringPointsDisplacementIsInput, "optimize measuring points", opt);
    flag = 0;
    ini.get(flag, "fast approximation is criterion", opt);
    fastApproxIsCriterion = flag > 0;
    fastApproxEpiDelay = 0;
    ini.get(fastApproxEpiDelay, "fast approximation epi delay", opt);
    fastApproxLimit = 2;
    ini.get(fastApproxLimit, "fast approximation limit", opt);
    int cmode = 2;
    ini.get(cmode, "comparison mode", opt);
    if (cmode < mode_rms || cmode > mode_vector_correlation) throw std::runtime_error("invalid comparison mode");
    comparisonMode = (ComparisonMode)cmode;
    numGenerations = 100;
    ini.get(numGenerations, "number of generations", opt);
    populationSize = 0;
    ini.get(populationSize, "population size", opt);
    queueSize = 1;
    ini.get(queueSize, "queue size", opt);
}

// ---- .column input, resampling -------------------------------------------------

std::vector<std::vector<double>> readColumns(const std::string& fname)
{
    ColumnFile f(fname.c_str());
    return f.columns();
}

void resample(const std::vector<double>& src, std::vector<double>& dest, double factor, int startOffset)
{
    int64_t n = 0;
    if (ekg_resample(src.data(), (int64_t)src.size(), factor, startOffset, nullptr, 0, &n) != EKG_OK)
        throw std::runtime_error(ekg_last_error());
    dest.assign((size_t)n, 0.0);
    if (ekg_resample(src.data(), (int64_t)src.size(), factor, startOffset, dest.data(), n, &n) != EKG_OK)
        throw std::runtime_error(ekg_last_error());
}

// ---- Evaluator ------------------------------------------------------------------

Evaluator::Evaluator(const char* ini, int device)
{
    std::cerr << "***** setting up Ekg Simulator *****************************\n";
    settings_.load(ini);
    sim_.reset(new SimLib::EkgSim(device));
    sim_->loadSettings(ini);

    sim_->loadTransferMatrix();
    sim_->loadMeasuringPoints();
    num_displacement_ = settings_.measuringPointsDisplacementIsInput ? sim_->numMeasurements() * 2 : 0;
    sim_->loadShape();
    sim_->simExcitationSequence();

    std::cerr << "\n***** applying (and checking) settings *********************\n";
    sim_->applySettings();
    if (settings_.interpolationTypeString == "endo-epi")
        interpolation_ = EKG_INTERP_ENDO_EPI;
    else if (settings_.interpolationTypeString == "endo-mid-epi")
        interpolation_ = EKG_INTERP_ENDO_MID_EPI;
    else
        throw std::runtime_error("unknown interpolation type [" + settings_.interpolationTypeString + "]");
    std::set<char> free_set;
    for (int k : settings_.freeKs) free_set.insert((char)k);
    num_wohlfart_ = free_set.size() * (interpolation_ == EKG_INTERP_ENDO_EPI ? 2 : 3);
    std::cerr << " ok\n";

    std::cerr << "\n***** printout of the simulator setup **********************\n";
    sim_->printSettings();

    std::cerr << "\n***** loading targets and setting up optimization **********\n";
    loadTargets(settings_.optimizationTargetsFname.c_str());

    size_t base = 0;
    if (settings_.criteriaMode == SimSettings::mode_every_lead_is_criterium)
        base = std::min(sim_->numMeasurements(), targets_.size());
    else if (settings_.criteriaMode == SimSettings::mode_leads_sum_is_criterium)
        base = 1;
    num_criteria_ = base;
    if (settings_.peakPositionIsCriterion) num_criteria_ += base;
    if (settings_.fastApproxIsCriterion) ++num_criteria_;
    if (settings_.endoEpiMinCriterionDelay >= 0) ++num_criteria_;
    std::cerr << " total number of criteria = " << num_criteria_ << "\n\n";
}

void Evaluator::loadTargets(const char* fname)
{
    // SimImplementation::loadTargets (sim.cpp:977-1042): the first column is
    // time when there are several; every target is scaled to unit range and
    // resampled to the simulation step starting at `start` -- except that a
    // factor of exactly 1 copies it unshifted.
    std::vector<std::vector<double>> file = readColumns(fname);
    const size_t ncol = file.size();
    targets_.assign(std::max<size_t>(1, ncol - 1), std::vector<double>());
    target_offsets_.assign(targets_.size(), 0.0);
    const SimLib::Settings& ss = sim_->getSettings();
    double step = ss.inputActionPotentialsTimeStep;
    const size_t time_cols = std::min<size_t>(1, ncol - 1);
    if (time_cols > 0) {
        if (file[0].size() < 2) throw std::runtime_error("target file needs at least two rows");
        step = file[0][1] - file[0][0];
    }
    const int start = (int)(ss.simulationStart / step);
    for (size_t i = time_cols; i < ncol; ++i) {
        std::vector<double>& col = file[i];
        double lo = col[0], hi = col[0];
        for (size_t j = 1; j < col.size(); ++j) {
            if (col[j] < lo)
                lo = col[j];
            else if (col[j] > hi)
                hi = col[j];
        }
        const double m = 1.0 / (hi - lo);
        for (double& v : col) v *= m;
        resample(col, targets_[i - time_cols], step / ss.simulationTimeStep, start);
        target_offsets_[i - time_cols] = 1.0 - (lo / (hi - lo));
    }
    std::vector<SimLib::ColumnVec> chk(targets_.size());
    for (size_t i = 0; i < chk.size(); ++i) {
        chk[i].name = "target " + std::to_string(i + 1);
        chk[i].data = &targets_[i];
    }
    SimLib::exportVectors(chk, "target_chk.column", ss.simulationStart, ss.simulationTimeStep);
}

double Evaluator::compare(const std::vector<double>& a, size_t t) const
{
    double v = 0;
    const std::vector<double>& b = targets_[t];
    if (ekg_criterion((int)settings_.comparisonMode, a.data(), (int64_t)a.size(), b.data(), (int64_t)b.size(),
                      target_offsets_[t], &v) != EKG_OK)
        throw std::runtime_error(ekg_last_error());
    return v;
}

void Evaluator::prepare(const std::vector<double>& x, Prepared& p)
{
    if (x.size() < num_wohlfart_ + num_displacement_)
        throw std::runtime_error("Solution does not contain enough values");
    // electrode displacement genes follow the AP genes (sim.cpp:456-468)
    p.points = sim_->originalMeasuringPoints();
    if (settings_.measuringPointsDisplacementIsInput)
        for (size_t i = 0; i < p.points.size(); ++i)
            ekg_move_lead(sim_->originalMeasuringPoints()[i].data(), x[num_wohlfart_ + 2 * i],
                          x[num_wohlfart_ + 2 * i + 1], p.points[i].data());
    else
        p.points = sim_->measuringPoints();
    if (x.size() != num_displacement_ + num_wohlfart_) {
        std::ostringstream s;
        s << "chromosome size does not agree with the combination of the number of free "
          << "Wohlfart parameters and the selected interpolation procedure (" << x.size()
          << " != " << num_displacement_ << "+" << num_wohlfart_ << ")";
        throw std::runtime_error(s.str());
    }
    // layer APs (sim.cpp:751-916)
    std::vector<double> base(9 * settings_.baseAps.size());
    for (size_t b = 0; b < settings_.baseAps.size(); ++b)
        std::copy(settings_.baseAps[b].begin(), settings_.baseAps[b].end(), base.begin() + 9 * b);
    ekg_ap_model m;
    m.n_base = (int)settings_.baseAps.size();
    m.base_k = base.data();
    m.k_min = settings_.kMin.data();
    m.k_max = settings_.kMax.data();
    m.n_free = (int)settings_.freeKs.size();
    m.free_k = settings_.freeKs.data();
    m.interpolation = interpolation_;
    m.mid_position = settings_.midPosition;
    const size_t nl = sim_->requiredAps();
    std::vector<double> k(9 * nl);
    if (ekg_layer_aps(&m, (int)nl, x.data(), (int)num_wohlfart_, k.data(), &p.violation) != EKG_OK)
        throw std::runtime_error(ekg_last_error());
    p.layerAps.assign(nl, SimLib::ActionPotential());
    for (size_t l = 0; l < nl; ++l) p.layerAps[l].init(&k[9 * l], 0);

    // fast approximation and the decision to simulate (sim.cpp:712-747)
    const SimLib::Settings& ss = sim_->getSettings();
    p.approx.assign((size_t)(ss.simulationLength / ss.simulationTimeStep), 0.0);
    for (size_t i = 0; i < p.approx.size(); ++i) {
        const double t = ss.simulationStart + i * ss.simulationTimeStep;
        p.approx[i] = p.layerAps[0](t + settings_.fastApproxEpiDelay) - p.layerAps.back()(t);
    }
    p.approxCriteria = 2;
    if (settings_.fastApproxIsCriterion || settings_.fastApproxLimit < 2) {
        p.approxCriteria = compare(p.approx, 0);  // against the first target
    }
    p.simulate = !(p.approxCriteria > settings_.fastApproxLimit);
}

void Evaluator::fitness(const Prepared& p, const std::vector<std::vector<double>>* ecg,
                        std::vector<double>& result, const double* crit) const
{
    // SimImplementation::calculateFitness (sim.cpp:600-702)
    if (settings_.criteriaMode != SimSettings::mode_every_lead_is_criterium)
        throw std::runtime_error("selected criteria mode is either not implemented yet or invalid");
    if (result.empty()) result.resize(num_criteria_);
    if (result.size() != num_criteria_)
        throw std::runtime_error("error in result.size() - doesn't match deducedNumOfCriteria");
    result[0] = 0;
    const size_t nm = std::min(sim_->numMeasurements(), targets_.size());
    if (settings_.fastApproxIsCriterion) result[nm] = p.approxCriteria;
    const bool dev = p.simulate && crit != nullptr;
    for (size_t i = 0; i < nm; ++i) {
        const std::vector<double>& sim = p.simulate && !dev ? (*ecg)[i] : p.approx;
        size_t ci = i;
        double value;
        if (dev) {
            // the device's values and peak distances (ekg_simulate_criteria)
            ci = settings_.peakPositionIsCriterion ? 2 * i : i;
            value = crit[ci];
            if (settings_.peakPositionIsCriterion) result[ci + 1] = crit[ci + 1];
        } else {
            value = p.simulate ? compare(sim, i) : p.approxCriteria;
            if (settings_.peakPositionIsCriterion) {
                ci = 2 * i;
                const long sp = (long)(std::max_element(sim.begin(), sim.end()) - sim.begin());
                const long tp =
                    (long)(std::max_element(targets_[i].begin(), targets_[i].end()) - targets_[i].begin());
                result[ci + 1] = (double)std::labs(sp - tp);
            }
        }
        result[ci] = value;
        if (result[ci] < 0 || !std::isfinite(result[ci]) || !std::isnormal(result[ci])) {
            std::cerr << " warning, ConvolutionResult returned wierd value: " << value << ", causing fitness to be "
                      << result[ci] << "; making correction - setting fitness to a large number\n";
            result[ci] = 2e10;
        }
    }
    if (settings_.endoEpiMinCriterionDelay >= 0) {
        double sd = 0.0;
        for (double t = 1.0; t < 700.0; t += 1.0) {
            const double d = p.layerAps.front()(t + settings_.endoEpiMinCriterionDelay) - p.layerAps.back()(t);
            sd += d * d;
        }
        result[result.size() - 1] = sd;
    }
}

namespace {

// OutputAps (sim.cpp:920-970): 700 samples at 1 ms per AP, comment
// "k = [k0, ..., k8]", column "ap_<index>".
void writeAps(const std::string& fname, const std::vector<SimLib::ActionPotential>& aps,
              const std::vector<size_t>& which)
{
    std::vector<std::vector<double>> data;
    std::vector<SimLib::ColumnVec> cols;
    data.reserve(which.size());
    for (size_t idx : which) {
        if (idx >= aps.size()) continue;
        const SimLib::ActionPotential& ap = aps[idx];
        std::vector<double> y(700);
        for (size_t t = 0; t < 700; ++t) y[t] = ap((double)t);
        data.push_back(y);
        SimLib::ColumnVec c;
        std::ostringstream k;
        k << "k = [" << ap[0];
        for (size_t i = 1; i < ap.size(); ++i) k << ", " << ap[i];
        k << "]";
        c.comment = k.str();
        c.name = "ap_" + std::to_string(idx);
        cols.push_back(c);
    }
    for (size_t i = 0; i < cols.size(); ++i) cols[i].data = &data[i];
    if (!cols.empty()) SimLib::exportVectors(cols, fname, 0, 1);
}

}  // namespace

void Evaluator::writeOutputs(const std::vector<double>& solution, const Prepared& p)
{
    if (!outSettings.outputCellAps.empty()) writeAps("cell_aps.column", sim_->getAps(), outSettings.outputCellAps);
    if (outSettings.layerAps) {
        std::vector<size_t> all(p.layerAps.size());
        for (size_t i = 0; i < all.size(); ++i) all[i] = i;
        writeAps("layer_aps.column", p.layerAps, all);
    }
    if (outSettings.result) sim_->saveMeasurement("input=" + angle(solution) + "; ");
}

double Evaluator::eval(const std::vector<double>& solution, std::vector<double>& result)
{
    ++evaluations_;
    std::cout << "\reval " << evaluations_ << "  ";
    Prepared p;
    prepare(solution, p);
    if (settings_.measuringPointsDisplacementIsInput) {
        std::vector<SimLib::PositionVec> d(sim_->numMeasurements());
        for (size_t i = 0; i < d.size(); ++i)
            d[i] = {solution[num_wohlfart_ + 2 * i], solution[num_wohlfart_ + 2 * i + 1], 0.0};
        sim_->moveMeasuringPoints(d);
    }
    std::vector<SimLib::ActionPotential> aps = p.layerAps;
    sim_->setApsDestructive(aps);
    std::vector<std::vector<double>> ecg;
    if (p.simulate) {
        sim_->run();
        for (size_t i = 0; i < sim_->numMeasurements(); ++i) ecg.push_back(sim_->getMeasurement(i));
    }
    double violation = p.violation;
    fitness(p, &ecg, result);
    writeOutputs(solution, p);
    return violation;
}

void Evaluator::evalBatch(const std::vector<std::vector<double>>& xs, std::vector<std::vector<double>>& results,
                          std::vector<double>& violations)
{
    const size_t B = xs.size();
    results.assign(B, std::vector<double>());
    violations.assign(B, 0.0);
    std::vector<Prepared> prep(B);
    for (size_t b = 0; b < B; ++b) prepare(xs[b], prep[b]);
    std::vector<size_t> run;
    for (size_t b = 0; b < B; ++b)
        if (prep[b].simulate) run.push_back(b);

    const SimLib::Settings& ss = sim_->getSettings();
    const size_t L = sim_->numMeasurements(), nl = sim_->requiredAps();
    const int64_t T = ekg_num_steps(ss.simulationLength, ss.simulationTimeStep);
    std::vector<std::vector<std::vector<double>>> ecg(B);
    // Criteria on the device when every target has the same length: only
    // the per-lead values and peak distances come back, not the ECGs.
    bool same_len = !targets_.empty();
    for (const auto& t : targets_) same_len = same_len && t.size() == targets_[0].size() && !t.empty();
    const size_t nm = std::min(L, targets_.size());
    const size_t per = nm * (settings_.peakPositionIsCriterion ? 2 : 1);
    std::vector<double> crit;
    std::vector<const double*> crit_of(B, nullptr);
    if (!run.empty()) {
        std::vector<double> k(run.size() * nl * 9), pos(run.size() * L * 3);
        std::vector<ekg_sim_desc> descs(run.size());
        for (size_t r = 0; r < run.size(); ++r) {
            const Prepared& p = prep[run[r]];
            for (size_t l = 0; l < nl; ++l) std::copy(p.layerAps[l].getK(), p.layerAps[l].getK() + 9, &k[(r * nl + l) * 9]);
            for (size_t i = 0; i < L; ++i)
                for (int c = 0; c < 3; ++c) pos[(r * L + i) * 3 + c] = p.points[i][c];
            ekg_sim_desc& d = descs[r];
            d.n_layers = (int)nl;
            d.layer_k = &k[r * nl * 9];
            d.n_leads = (int)L;
            d.lead_pos = &pos[r * L * 3];
            d.nhood = sim_->neighbourhood();
            d.start = ss.simulationStart;
            d.length = ss.simulationLength;
            d.dt = ss.simulationTimeStep;
        }
        if (same_len && nm > 0 && T > 0) {
            std::vector<double> tflat;
            for (const auto& t : targets_) tflat.insert(tflat.end(), t.begin(), t.end());
            ekg_criteria_desc c{};
            c.mode = (int)settings_.comparisonMode;
            c.peaks = settings_.peakPositionIsCriterion ? 1 : 0;
            c.raw = 1;  // the correction and its warning follow in fitness()
            c.n_targets = (int)targets_.size();
            c.target_len = (int64_t)targets_[0].size();
            c.targets = tflat.data();
            c.target_offsets = target_offsets_.data();
            crit.assign(run.size() * per, 0.0);
            if (ekg_simulate_criteria(sim_->heart(), (int)run.size(), descs.data(), &c, crit.data()) != EKG_OK)
                throw std::runtime_error(ekg_last_error());
            for (size_t r = 0; r < run.size(); ++r) crit_of[run[r]] = crit.data() + r * per;
        } else {
            std::vector<double> out(run.size() * L * (size_t)T);
            if (ekg_simulate_batch(sim_->heart(), (int)run.size(), descs.data(), out.data()) != EKG_OK)
                throw std::runtime_error(ekg_last_error());
            for (size_t r = 0; r < run.size(); ++r) {
                auto& e = ecg[run[r]];
                e.resize(L);
                for (size_t i = 0; i < L; ++i)
                    e[i].assign(out.begin() + (r * L + i) * (size_t)T, out.begin() + (r * L + i + 1) * (size_t)T);
            }
        }
    }
    for (size_t b = 0; b < B; ++b) {
        fitness(prep[b], &ecg[b], results[b], crit_of[b]);
        violations[b] = prep[b].violation;
    }
    evaluations_ += B;
}

void Evaluator::getGeneParams(size_t& numGenes, size_t& numCriteria, std::vector<double>& gMin,
                              std::vector<double>& gMax)
{
    // sim.cpp:512-578: per free coefficient its [k min, k max], repeated per
    // border/mid AP, then the displacement bounds (last value repeats).
    std::set<char> free_set;
    for (int k : settings_.freeKs) free_set.insert((char)k);
    const size_t reps = interpolation_ == EKG_INTERP_ENDO_EPI ? 2 : 3;
    gMin.assign(free_set.size() * reps, 0.0);
    gMax.assign(free_set.size() * reps, 0.0);
    size_t g = 0;
    for (int i = 0; i < 9; ++i)
        if (free_set.count((char)i)) {
            for (size_t r = 0; r < reps; ++r) {
                gMin[g + r * free_set.size()] = settings_.kMin[i];
                gMax[g + r * free_set.size()] = settings_.kMax[i];
            }
            ++g;
        }
    numGenes = free_set.size() * reps;
    numCriteria = num_criteria_;
    if (settings_.measuringPointsDisplacementIsInput) {
        numGenes += num_displacement_;
        std::vector<int> lo = settings_.displacementMin, hi = settings_.displacementMax;
        if (lo.empty()) lo.push_back(0);
        if (hi.empty()) hi.push_back(0);
        for (size_t i = 0; i < num_displacement_; ++i) {
            gMin.push_back(lo[std::min(i, lo.size() - 1)]);
            gMax.push_back(hi[std::min(i, hi.size() - 1)]);
        }
    }
}

}  // namespace ekgapp
