// sim_lib.cpp -- SimLib::EkgSim facade over the GPU C ABI; see sim_lib.hpp.
#include "sim_lib.hpp"

#include <cmath>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <sstream>
#include <stdexcept>

#include "ini.hpp"

namespace SimLib {

namespace {

std::string extension(const std::string& f)
{
    const size_t p = f.rfind('.');
    return p == std::string::npos ? std::string() : f.substr(p);
}

// Header of a .matrix file (matrix.h:139-183): a comment line, "<n>D" with
// n <= dims, the sizes "X x Y [x Z]" filled from the last dimension backwards,
// then an optional separator word (tab, comma or ",", space; anything else
// means one character per element).
struct MatrixHeader {
    size_t size[3] = {1, 1, 1};  // slowest first; only the last `dims` are used
    char sep = 0;
};

MatrixHeader read_header(std::istream& in, int dims, const std::string& fname)
{
    MatrixHeader h;
    in.ignore(100000000, '\n');
    std::string dim;
    in >> dim;
    if (!(dim.size() >= 2 && (dim[1] == 'D' || dim[1] == 'd') && dim[0] > '0' && dim[0] <= '0' + dims))
        throw std::runtime_error("error while loading matrix, could not read dimensionality");
    int n = dim[0] - '0';
    int w = dims - 1;
    in >> h.size[w];
    for (--n; n > 0; --n) {
        in.ignore(100, 'x');
        --w;
        in >> h.size[w];
    }
    std::string rest, word;
    std::getline(in, rest, '\n');
    std::istringstream rs(rest);
    rs >> word;
    h.sep = word == "tab" ? '\t' : (word == "comma" || word == ",") ? ',' : word == "space" ? ' ' : 0;
    if (in.fail()) throw std::runtime_error("reading file " + fname + " failed (while reading size)");
    return h;
}

// Element stream of importMatrix (matrix.h:185-240): numbers between the
// declared separators, or single characters ('x'/'X' = seed of layer 1,
// letters = 10 + offset, digits) when none is declared.
template <class Load>
void read_elements(std::istream& in, const MatrixHeader& h, int dims, size_t count, Load load,
                   const std::string& fname)
{
    const size_t row = h.size[dims - 1];
    for (size_t i = 0; i < count; ++i) {
        if (h.sep == ',' || h.sep == '\t' || h.sep == ' ') {
            load(i, in, false, 0);
        } else {
            char ch = 0;
            in >> ch;
            load(i, in, true, ch);
        }
        const bool row_end = ((i + 1) % row) == 0;
        if (h.sep == ',' && !row_end) in.ignore(100, ',');
    }
    if (in.fail()) throw std::runtime_error("reading file " + fname + " failed (while reading data)");
}

double char_value(char ch)
{
    if (ch == 'x' || ch == 'X') return -1.0;  // seed, layer 1
    if (ch >= 'a') return ch - 'a' + 10;
    if (ch >= 'A') return ch - 'A' + 10;
    return ch - '0';
}

}  // namespace

// ---- ActionPotential -------------------------------------------------------

void ActionPotential::init(const double* wohlK, double newAt)
{
    // ActionPotential::init (simulator.cpp:154-160): k8 moves with the
    // activation time, so the curve is read at t - at
    double* k = wohl.getK();
    for (int i = 0; i < 9; ++i) k[i] = wohlK[i];
    k[8] -= newAt;
    at = newAt;
}

void ActionPotential::init(const ActionPotential& ap, double newAt) { init(ap.getK(), newAt); }

// ---- Settings --------------------------------------------------------------

void Settings::loadFromIni(const ekgapp::IniFile& ini)
{
    const int model = ini.section("model");
    ini.get(inputShapeFilename, "shape", model);
    ini.get(inputExcitationSequenceFilename, "excitation sequence", model);
    ini.get(inputPointsFilename, "points", model);
    ini.get(inputTransferFilename, "transfer", model);

    const int aps = ini.section("action potentials");
    ini.get(inputActionPotentials, "input file", aps);
    ini.get(inputActionPotentialsFilter, "filter", aps);
    ini.get(inputActionPotentialsTimeStep, "resample time step", aps);
    ini.get(inputActionPotentialsFilterParam, "filter parameter", aps);
    ini.get(inputActionPotentialsFunction, "combination function", aps);
    ini.get(inputActionPotentialsScale, "expected length", aps);

    const int out = ini.section("output files");
    ini.get(outputExcitationSequence, "excitation sequence", out);
    ini.get(outputApFilename, "action potentials filename", out);
    outputFilename = "result.column";
    ini.get(outputFilename, "results filename", out);

    const int sim = ini.section("simulation");
    ini.get(simulationLength, "length", sim);
    ini.get(simulationStart, "start", sim);
    ini.get(neighbourhoodType, "neighbourhood type", sim);
    simulationTimeStep = 1.0 / 3.0;
    ini.get(simulationTimeStep, "time step", sim);
}

// ---- file formats ----------------------------------------------------------

void importShape(const std::string& fname, ModelShape& s)
{
    if (extension(fname) != ".matrix")
        throw std::runtime_error("can only import from .matrix files for the time being");
    std::ifstream in(fname.c_str());
    if (!in.is_open()) throw std::runtime_error("could not open file " + fname);
    const MatrixHeader h = read_header(in, 3, fname);
    for (int d = 0; d < 3; ++d) {
        if (h.size[d] < 1 || h.size[d] > 65535) throw std::runtime_error("unsupported matrix size in " + fname);
        s.size[d] = (int)h.size[d];
    }
    const size_t n = s.voxels();
    s.layer.assign(n, 0);
    s.seed.assign(n, 0);
    s.activation.assign(n, 0.0);
    read_elements(in, h, 3, n,
                  [&](size_t i, std::istream& is, bool chars, char ch) {
                      long v;
                      if (chars) {
                          v = (long)char_value(ch);
                      } else {
                          int x = 0;
                          is >> x;
                          v = x;
                      }
                      // negative = excitation start of layer -v (matrix.h:124-129)
                      const long layer = v < 0 ? -v : v;
                      if (layer > 255) throw std::runtime_error("layer numbers above 255 are not supported");
                      s.layer[i] = (uint8_t)layer;
                      s.seed[i] = v < 0 ? 1 : 0;
                  },
                  fname);
}

void importTransfer(const std::string& fname, std::vector<double>& t, int& rows, int& cols)
{
    if (extension(fname) != ".matrix")
        throw std::runtime_error("can only import from .matrix files for the time being");
    std::ifstream in(fname.c_str());
    if (!in.is_open()) throw std::runtime_error("could not open file " + fname);
    // a 2-D matrix: T[from][to], the second size (columns) read first
    const MatrixHeader h = read_header(in, 2, fname);
    if (h.size[0] > 4096 || h.size[1] > 4096) throw std::runtime_error("transfer matrix too large in " + fname);
    rows = (int)h.size[0];
    cols = (int)h.size[1];
    t.assign((size_t)rows * cols, 0.0);
    read_elements(in, h, 2, t.size(),
                  [&](size_t i, std::istream& is, bool chars, char ch) {
                      if (chars) {
                          const double v = char_value(ch);
                          t[i] = v < 0 ? 4097.0 : v;  // the seed flag value 0x1000 + 1
                      } else {
                          is >> t[i];
                      }
                  },
                  fname);
}

void importExcitation(const std::string& fname, ModelShape& s)
{
    // Simulation::loadExcitationSequence (simulator.cpp:288-367)
    if (extension(fname) != ".matrix") throw std::runtime_error("can only import from .matrix for the time being");
    std::ifstream in(fname.c_str());
    if (!in.is_open()) throw std::runtime_error("could not open file " + fname + " to load excitation sequence");
    const MatrixHeader h = read_header(in, 3, fname);
    if ((int)h.size[0] != s.size[0] || (int)h.size[1] != s.size[1] || (int)h.size[2] != s.size[2])
        throw std::runtime_error("failed to read excitation sequence due to its incompatible size");
    const size_t n = s.voxels();
    s.activation.resize(n, 0.0);
    for (size_t i = 0; i < n; ++i) {
        s.seed[i] = 0;  // the start flag is cleared on load
        in >> s.activation[i];
        if (h.sep == ',') in.ignore(1);
    }
}

void exportShape(const ModelShape& s, const std::string& fname, bool excit)
{
    // exportShape (simulator.cpp:71-106)
    if (extension(fname) != ".matrix") throw std::runtime_error("can only export to .matrix for the time being");
    std::ofstream out(fname.c_str());
    out << (excit ? "Excitation" : "Shape")
        << " file (next line specifies matrix type (2D) size (X x Y x Z) and elemet separator type (tab) after "
           "that comes data (x are rows, y are lines, z are blocks - separated with empty line)\n";
    if (s.size[0] > 1)
        out << "3D " << s.size[2] << " x " << s.size[1] << " x " << s.size[0];
    else
        out << "2D " << s.size[2] << " x " << s.size[1];
    out << (excit ? " tab\n" : "\n");
    size_t i = 0;
    for (int z = 0; z < s.size[0]; ++z) {
        for (int y = 0; y < s.size[1]; ++y) {
            for (int x = 0; x < s.size[2]; ++x, ++i) {
                if (excit) {
                    if (x > 0) out << '\t';
                    out << std::fixed << std::setprecision(3) << s.activation[i];
                } else {
                    const int l = s.layer[i];
                    out << (s.seed[i] ? 'X' : (l > 9 ? (char)(l - 10 + 'A') : (char)('0' + l)));
                }
            }
            out << "\n";
        }
        out << "\n";
    }
}

void generateTestShape(ModelShape& s)
{
    // InputLoader::generateTestShape (simulator.h:600-644): a 2-D ring of
    // 12 layers, 2 voxels each, on a 120 x 160 plane.
    s.size[0] = 1;
    s.size[1] = 160;
    s.size[2] = 120;
    const size_t n = s.voxels();
    s.layer.assign(n, 0);
    s.seed.assign(n, 0);
    s.activation.assign(n, 0.0);
    const long mid1 = 160 / 2, mid2 = 120 / 4;
    for (long i = 0; i < 160; ++i)
        for (long j = 0; j < 120; ++j) {
            const double d = std::sqrt((double)((i - mid1) * (i - mid1)) + (double)((j - mid2) * (j - mid2)));
            if (d > 50 && d < 76) s.layer[(size_t)i * 120 + j] = (uint8_t)((size_t)(d - 50) / 2);
        }
    long y = 160 / 2;
    for (; y < 160; ++y)
        if (s.layer[(size_t)y * 120 + 30] > 0) break;
    if (y == 160) throw std::runtime_error("Could not create starting point for excitation sequence");
    s.seed[(size_t)y * 120 + 30] = 1;
    exportShape(s, "autoGeneratedShape.matrix", false);
}

void exportVectors(const std::vector<saveVecElement>& vec, const std::string& fname, double startTime, double timeStep,
                   const std::string& comment)
{
    // exportVectors (simulator.h:194-236); the stream keeps the precision
    // set for the time column, so values print with 5 digits after a time
    // column has been written.
    if (extension(fname) != ".column")
        throw std::runtime_error("saving vectors to .column files only for the time being");
    std::ofstream out(fname.c_str());
    out << "#Comment: " << comment << '\n';
    for (const saveVecElement& v : vec)
        if (!v.comment.empty()) out << "# " << v.name << " : " << v.comment << '\n';
    out << "# \n";
    out << (timeStep > 0 ? "#Time[ms]\t" : "# ");
    for (size_t i = 0; i < vec.size(); ++i) {
        if (i > 0) out << '\t';
        out << vec[i].name;
    }
    out << '\n';
    if (vec.empty()) return;
    double time = startTime;
    for (size_t i = 0; i < vec[0].data->size(); ++i) {
        if (i > 0) out << "\n";
        if (timeStep > 0) {
            out << std::fixed << std::setprecision(5) << std::setw(10) << time << '\t';
            time += timeStep;
        }
        for (size_t j = 0; j < vec.size(); ++j) {
            if (j > 0) out << '\t';
            out << std::scientific << std::setw(10) << (*vec[j].data)[i];
        }
    }
}

// ---- EkgSim ----------------------------------------------------------------

EkgSim::EkgSim(int device)
{
    if (device < 0) {
        const char* e = std::getenv("EKG_DEVICE");
        device = e ? std::atoi(e) : 0;
    }
    check(ekg_ctx_create(device, &ctx_), "GPU context");
}

EkgSim::~EkgSim()
{
    if (heart_) ekg_heart_destroy(heart_);
    if (ctx_) ekg_ctx_destroy(ctx_);
}

void EkgSim::check(int rc, const char* what) const
{
    if (rc == EKG_OK) return;
    if (rc == EKG_ENOSEED) throw std::runtime_error("Could not find starting point for excitation sequence");
    throw std::runtime_error(std::string(what) + ": " + ekg_last_error());
}

void EkgSim::loadSettings(const char* fname)
{
    ekgapp::IniFile ini(fname);
    if (!ini.found())
        std::cout << "warning, " << fname << " not found. ";
    else
        settings_.loadFromIni(ini);
}

void EkgSim::applySettings()
{
    const std::string& t = settings_.neighbourhoodType;
    if (t == "2D4")
        nhood_ = EKG_NHOOD_2D4;
    else if (t == "2D8")
        nhood_ = EKG_NHOOD_2D8;
    else if (t == "3D4")
        nhood_ = EKG_NHOOD_3D4;
    else if (t == "3D8" || t == "cube")
        nhood_ = EKG_NHOOD_3D8;
    else
        throw std::runtime_error("\n   unknown neighbourhood (only know of these: 2D4, 3D4, 2D8, 3D8 (cube))");
    time_step_ = settings_.simulationTimeStep;
}

void EkgSim::setApsDestructive(std::vector<ActionPotential>& destructible)
{
    aps_.swap(destructible);
    aps_are_cells_ = false;
    cells_built_ = false;
}

const std::vector<ActionPotential>& EkgSim::getAps() const
{
    if (aps_are_cells_ && !cells_built_) buildCellAps();
    return aps_;
}

void EkgSim::buildCellAps() const
{
    // Simulation::setApIndices (simulator.cpp:561-621): one AP per distinct
    // (layer, activation time), numbered in raster order of first use.
    std::vector<std::map<double, size_t>> index(layer_aps_.size());
    std::vector<std::pair<int, double>> classes;
    for (size_t i = 0; i < shape_.voxels(); ++i) {
        const int l = shape_.layer[i];
        if (l == 0) continue;
        auto& m = index[l - 1];
        if (m.find(shape_.activation[i]) == m.end()) {
            m[shape_.activation[i]] = classes.size();
            classes.emplace_back(l, shape_.activation[i]);
        }
    }
    aps_.assign(classes.size(), ActionPotential());
    for (size_t c = 0; c < classes.size(); ++c) aps_[c].init(layer_aps_[classes[c].first - 1], classes[c].second);
    cells_built_ = true;
}

void EkgSim::loadTransferMatrix()
{
    importTransfer(settings_.inputTransferFilename, transfer_, transfer_rows_, transfer_cols_);
    if ((size_t)transfer_rows_ < target_aps_ || (size_t)transfer_cols_ < target_aps_)
        throw std::runtime_error("loaded transfer matrix too small");
}

void EkgSim::loadMeasuringPoints()
{
    // Simulation::loadMeasuringPoints (simulator.cpp:369-396): lines
    // "x, y, z" stored as (z, y, x); reading stops at the first bad line.
    const std::string& fname = settings_.inputPointsFilename;
    std::ifstream in(fname.c_str());
    std::vector<PositionVec> pts;
    if (in.is_open()) {
        while (!in.eof()) {
            PositionVec p{};
            in >> p[2];
            in.ignore(1000, ',');
            in >> p[1];
            in.ignore(1000, ',');
            in >> p[0];
            in.ignore(1000, '\n');
            if (in.fail()) break;
            pts.push_back(p);
        }
        original_ = pts;
    }
    for (const PositionVec& p : original_) std::cerr << " -> " << p[2] << ", " << p[1] << ", " << p[0] << "\n";
    if (original_.empty()) throw std::runtime_error("failed to open file containing measuring points: " + fname);
    if (original_.size() > EKG_MAX_LEADS)
        throw std::runtime_error("more than " + std::to_string(EKG_MAX_LEADS) + " measuring points");
    // U/V displacement planes (EkgSim::loadMeasuringPoints, sim_lib.h:167-193)
    points_ = original_;
    u_.resize(original_.size());
    v_.resize(original_.size());
    for (size_t i = 0; i < original_.size(); ++i) {
        ekg_lead_basis(original_[i].data(), u_[i].data(), v_[i].data());
        std::cout << " u" << i << " = " << u_[i][2] << "," << u_[i][1] << "," << u_[i][0] << "\n";
        std::cout << " v" << i << " = " << v_[i][2] << "," << v_[i][1] << "," << v_[i][0] << "\n";
    }
}

void EkgSim::moveMeasuringPoints(const std::vector<PositionVec>& d)
{
    if (d.size() != points_.size()) {
        std::cerr << d.size() << " != " << points_.size() << " " << original_.size() << std::endl;
        throw std::runtime_error("invalid vector of measuring point displacements");
    }
    for (size_t i = 0; i < original_.size(); ++i) ekg_move_lead(original_[i].data(), d[i][0], d[i][1], points_[i].data());
}

void EkgSim::loadShape()
{
    if (settings_.inputShapeFilename.empty())
        generateTestShape(shape_);
    else
        importShape(settings_.inputShapeFilename, shape_);
    size_t max_layer = 0;
    for (uint8_t l : shape_.layer) max_layer = std::max<size_t>(max_layer, l);
    target_aps_ = max_layer;
    if (heart_) ekg_heart_destroy(heart_);
    heart_ = nullptr;
    excited_ = false;
    check(ekg_heart_create(ctx_, shape_.size[0], shape_.size[1], shape_.size[2], shape_.layer.data(),
                           shape_.seed.data(), &heart_),
          "heart upload");
}

void EkgSim::simExcitationSequence()
{
    if (!heart_) throw std::runtime_error("no heart model loaded");
    if (settings_.inputExcitationSequenceFilename.empty()) {
        // Simulation::calculateExcitationSequence (simulator.cpp:248-286) on the GPU
        if (transfer_.empty())
            throw std::runtime_error("transfer (conduction) matrix does not define layer " + std::to_string(target_aps_));
        check(ekg_heart_excite_transfer(heart_, transfer_.data(), transfer_rows_, transfer_cols_), "excitation");
        shape_.activation.resize(shape_.voxels());
        check(ekg_heart_get_activation(heart_, shape_.activation.data()), "activation download");
        std::fill(shape_.seed.begin(), shape_.seed.end(), 0);
    } else {
        importExcitation(settings_.inputExcitationSequenceFilename, shape_);
        check(ekg_heart_set_activation(heart_, shape_.activation.data()), "activation upload");
    }
    excited_ = true;
    if (!settings_.outputExcitationSequence.empty()) exportShape(shape_, settings_.outputExcitationSequence, true);
}

void EkgSim::printSettings()
{
    const Settings& s = settings_;
    std::cout << "model: " << s.inputShapeFilename << " (" << shape_.size[0] << ", " << shape_.size[1] << ", "
              << shape_.size[2] << ")\n";
    std::cout << "neighbourhood: ";
    if (s.neighbourhoodType == "2D4")
        std::cout << "2D, 4 neighbours\n";
    else if (s.neighbourhoodType == "2D8")
        std::cout << "2D, 8 neighbours\n";
    else if (s.neighbourhoodType == "3D4")
        std::cout << "3D, 8 corner neighbours\n";
    else if (s.neighbourhoodType == "3D8" || s.neighbourhoodType == "cube")
        std::cout << "3D, 26 neighbours (full cube)\n";
    std::cout << "simulation start = " << s.simulationStart << "\n"
              << "simulation time step = " << s.simulationTimeStep << "\n"
              << "simulation length = " << s.simulationLength << "\n"
              << "total steps = " << s.simulationLength / s.simulationTimeStep << "\n";
}

std::vector<double> EkgSim::layerCoefficients() const
{
    const std::vector<ActionPotential>& src = aps_are_cells_ ? layer_aps_ : aps_;
    std::vector<double> k(9 * src.size());
    for (size_t l = 0; l < src.size(); ++l)
        for (int i = 0; i < 9; ++i) k[9 * l + i] = src[l][i];
    return k;
}

void EkgSim::run()
{
    // Simulation::run (simulator.cpp:452-550) on the GPU.
    if (!excited_) throw std::runtime_error("excitation sequence missing: call simExcitationSequence() first");
    if (nhood_ < 0) throw std::runtime_error("neighbourhood not set: call applySettings() first");
    if (aps_are_cells_) {
        // a second run() without new APs: the reference would rebuild cell
        // classes from cell classes; keep the layer APs of the last call
    } else {
        if (aps_.size() != target_aps_)
            throw std::runtime_error("number of APs (" + std::to_string(aps_.size()) +
                                     ") differs from the number of layers (" + std::to_string(target_aps_) + ")");
        layer_aps_ = aps_;
    }
    const std::vector<double> k = layerCoefficients();
    std::vector<double> pos(3 * points_.size());
    for (size_t i = 0; i < points_.size(); ++i)
        for (int c = 0; c < 3; ++c) pos[3 * i + c] = points_[i][c];
    ekg_sim_desc d;
    d.n_layers = (int)layer_aps_.size();
    d.layer_k = k.data();
    d.n_leads = (int)points_.size();
    d.lead_pos = pos.data();
    d.nhood = nhood_;
    d.start = settings_.simulationStart;
    d.length = settings_.simulationLength;
    d.dt = time_step_;
    const int64_t T = ekg_num_steps(d.length, d.dt);
    if (T < 0) throw std::runtime_error("invalid simulation length or time step");
    std::vector<double> ecg((size_t)T * points_.size());
    check(ekg_simulate(heart_, &d, ecg.data()), "simulation");
    measurements_.assign(points_.size(), std::vector<double>());
    for (size_t i = 0; i < points_.size(); ++i)
        measurements_[i].assign(ecg.begin() + (size_t)T * i, ecg.begin() + (size_t)T * (i + 1));
    start_time_ = settings_.simulationStart;
    aps_are_cells_ = true;
    cells_built_ = false;
}

void EkgSim::runApproximation(double delay, std::vector<double>& result)
{
    // Simulation::runApproximation (simulator.cpp:552-559): endo AP shifted by
    // `delay` minus the last AP, t = start + i dt.
    const std::vector<ActionPotential>& aps = getAps();
    if (aps.empty()) throw std::runtime_error("no action potentials set");
    result.resize((size_t)(settings_.simulationLength / time_step_), 0.0);
    for (size_t i = 0; i < result.size(); ++i) {
        const double t = settings_.simulationStart + i * time_step_;
        result[i] = aps[0](t + delay) - aps.back()(t);
    }
}

const std::vector<double>& EkgSim::getMeasurement(size_t n) const { return measurements_.at(n); }

void EkgSim::saveTo(const std::string& fname, const std::string& comment)
{
    // Simulation::saveMeasurements (simulator.cpp:413-427): one column per
    // measuring point, named "x,y,z".
    if (points_.empty()) return;
    std::vector<ColumnVec> cols(points_.size());
    static const std::vector<double> none;
    for (size_t i = 0; i < points_.size(); ++i) {
        std::ostringstream name;
        name << points_[i][2] << "," << points_[i][1] << "," << points_[i][0];
        cols[i].name = name.str();
        cols[i].data = i < measurements_.size() ? &measurements_[i] : &none;
    }
    exportVectors(cols, fname, start_time_, time_step_, comment);
}

void EkgSim::saveMeasurement(const std::string& comment) { saveTo(settings_.outputFilename, comment); }

void EkgSim::saveMeasurementAs(const char* fname, const std::string& comment) { saveTo(fname, comment); }

}  // namespace SimLib

// ---- ColumnFile --------------------------------------------------------------

void ColumnFile::load(const char* fname)
{
    // ColumnFile::load (columnFile.h:124-170): skip '#'/'%' lines, the first
    // data line fixes the column count, then one row per line until a line
    // holds fewer numbers or the file ends
    const std::string f(fname);
    const size_t dot = f.rfind('.');
    if (dot == std::string::npos || f.substr(dot) != ".column")
        throw std::runtime_error("ColumnFile::load only works on .column files");
    std::ifstream in(fname);
    if (!in.is_open()) throw std::runtime_error("could not open " + f);
    while (in.peek() == '#' || in.peek() == '%') in.ignore(10000, '\n');
    std::string line;
    std::getline(in, line);
    std::vector<double> row;
    {
        std::istringstream ls(line);
        double d;
        while (ls >> d) row.push_back(d);
    }
    if (row.empty()) throw std::runtime_error("File has 0 columns");
    cols_.assign(row.size(), Column());
    bool ok = (bool)in;
    while (ok) {
        for (size_t i = 0; i < cols_.size(); ++i) cols_[i].push_back(row[i]);
        std::getline(in, line);
        if (!in) break;
        std::istringstream ls(line);
        for (size_t i = 0; i < row.size(); ++i) ls >> row[i];
        ok = (bool)ls;
    }
}
