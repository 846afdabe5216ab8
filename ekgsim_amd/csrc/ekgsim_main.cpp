// ekgsim_main.cpp -- the ekgSim command line (main.cpp:150-415) on the GPU.
//
//   ekgSim [free] -sim x0,x1,...  [-out result] [-out layer_aps] [-out "cell_aps 3 7"]
//       one evaluation from ./simulator.ini: writes the requested files and
//       prints " criteria = <c1,...>, violation = v"
//   ekgSim -batch inputs.txt [-batch-out outputs.txt]
//       many vectors (one per line, ',' ';' or blank separated) in one batched
//       GPU simulation; one line "criteria... violation" per vector
//   ekgSim -external <dir>
//       one evaluation in the file protocol of AMS-DEMO's ExternalEvaluation
//       (ExternalEvaluation.h:95-151): <dir>/input.txt holds one value per
//       line after a '#' comment line; <dir>/output.txt receives
//       "# violation v" and one criterion per line
//
// Arguments follow the reference's DashArguments rules (Arguments.cpp): a
// word starting with '-' and not a digit is a key, the next word is its value,
// further words are free arguments; keys may repeat (-out).
#include <cctype>
#include <chrono>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "ekgsim_app.hpp"

namespace {

struct DashArgs {
    std::multimap<std::string, std::string> keys;
    std::vector<std::string> free;

    DashArgs(int argc, char** argv)
    {
        auto last = keys.end();
        for (int i = 1; i < argc; ++i) {
            const std::string a = argv[i];
            if (a.size() > 1 && a[0] == '-' && !std::isdigit((unsigned char)a[1])) {
                last = keys.insert(std::make_pair(a, std::string()));
            } else if (!keys.empty() && last->second.empty()) {
                last->second = a;
            } else {
                free.push_back(a);
            }
        }
    }
    bool has(const std::string& k) const { return keys.count(k) > 0; }
    std::string operator[](const std::string& k) const
    {
        auto it = keys.find(k);
        return it == keys.end() ? std::string() : it->second;
    }
    std::vector<std::string> all(const std::string& k) const
    {
        std::vector<std::string> v;
        auto r = keys.equal_range(k);
        for (auto it = r.first; it != r.second; ++it) v.push_back(it->second);
        return v;
    }
};

std::string angle(const std::vector<double>& v)
{
    std::ostringstream s;
    s << "<";
    for (size_t i = 0; i < v.size(); ++i) s << (i ? "," : "") << v[i];
    s << ">";
    return s.str();
}

// "-sim" value: numbers separated by single ',' or ';' (main.cpp:213-219).
std::vector<double> parse_vector(const std::string& text)
{
    std::istringstream in(text);
    std::vector<double> v;
    for (double x; in >> x;) {
        v.push_back(x);
        const int ch = in.peek();
        if (ch == ',' || ch == ';') in.ignore(1);
    }
    return v;
}

// "-out" values (main.cpp:244-272).
void parse_outputs(const std::vector<std::string>& vals, ekgapp::OutputSettings& out)
{
    for (const std::string& v : vals) {
        std::istringstream s(v);
        std::string name;
        s >> name;
        if (name == "cell_aps") {
            while (s) {
                s.ignore(1);
                size_t n;
                s >> n;
                if (s) out.outputCellAps.push_back(n);
            }
        } else if (name == "layer_aps") {
            out.layerAps = true;
        } else if (name == "result") {
            out.result = true;
        } else {
            std::cerr << "skipping an unrecognized output opition: " << name << "\n";
        }
    }
}

void show_help()
{
    std::cout << "Argument list:\n"
              << "   -?     \tshow this help screen\n"
              << "   -batch \tevaluate every vector of a file (one per line) in one batched GPU run\n"
              << "   -batch-out \tfile for the -batch results (default: stdout)\n"
              << "   -external \tevaluate <dir>/input.txt into <dir>/output.txt (ExternalEvaluation protocol)\n"
              << "   -out   \tspecify outputs of the program; possible values include result, layer_aps, "
                 "cell_aps <num> [<num>]*; multiple outputs are allowed, e.g. -out result -out layer_aps\n"
              << "   -sim   \tjust run single a simulation with parameters provided after -sim\n";
}

int run_sim(const DashArgs& args)
{
    std::vector<double> x = parse_vector(args["-sim"]);
    std::cout << " parameters for single simulator run: " << angle(x) << "\n";
    if (x.empty()) throw std::runtime_error(" Error: not enough parameters to run simulation: " + args["-sim"]);
    ekgapp::OutputSettings out;
    parse_outputs(args.all("-out"), out);
    std::cerr << "##### Running a single simulation experiment ####################\n";
    ekgapp::Evaluator ev("simulator.ini");
    ev.outSettings = out;
    std::vector<double> result;
    const auto t0 = std::chrono::steady_clock::now();
    const double violation = ev.eval(x, result);
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::cout << " simulation done in " << s << " seconds\n";
    std::cout << " criteria = " << angle(result) << ", violation = " << violation << "\n";
    return 0;
}

int run_batch(const DashArgs& args)
{
    std::ifstream in(args["-batch"].c_str());
    if (!in.is_open()) throw std::runtime_error("could not open " + args["-batch"]);
    std::vector<std::vector<double>> xs;
    for (std::string line; std::getline(in, line);) {
        if (line.empty() || line[0] == '#') continue;
        for (char& c : line)
            if (c == ',' || c == ';' || c == '\t') c = ' ';
        std::istringstream ls(line);
        std::vector<double> v;
        for (double d; ls >> d;) v.push_back(d);
        if (!v.empty()) xs.push_back(v);
    }
    ekgapp::Evaluator ev("simulator.ini");
    std::vector<std::vector<double>> results;
    std::vector<double> viol;
    const auto t0 = std::chrono::steady_clock::now();
    ev.evalBatch(xs, results, viol);
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::ofstream file;
    const std::string dest = args["-batch-out"];
    if (!dest.empty()) file.open(dest.c_str());
    std::ostream& o = dest.empty() ? std::cout : file;
    o.precision(17);
    for (size_t b = 0; b < xs.size(); ++b) {
        for (double c : results[b]) o << c << " ";
        o << viol[b] << "\n";
    }
    std::cerr << " " << xs.size() << " evaluations in " << s << " seconds\n";
    return 0;
}

int run_external(const DashArgs& args)
{
    const std::string dir = args["-external"];
    const std::string inName = dir.empty() ? "input.txt" : dir + "/input.txt";
    const std::string outName = dir.empty() ? "output.txt" : dir + "/output.txt";
    std::ifstream in(inName.c_str());
    if (!in.is_open()) throw std::runtime_error("could not open " + inName);
    std::vector<double> x;
    for (std::string line; std::getline(in, line);) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream ls(line);
        for (double d; ls >> d;) x.push_back(d);
    }
    ekgapp::Evaluator ev("simulator.ini");
    std::vector<double> result;
    const double violation = ev.eval(x, result);
    std::ofstream out(outName.c_str());
    out.precision(17);
    out << "# violation " << violation << "\n";
    for (double c : result) out << c << "\n";
    if (!out) throw std::runtime_error("writing " + outName + " failed");
    return 0;
}

}  // namespace

int main(int argc, char** argv)
{
    try {
        DashArgs args(argc, argv);
        if (args.has("-?")) {
            show_help();
            return 0;
        }
        if (args.has("-test")) {
            std::cout << "the -test function of the reference is not part of this build\n";
        } else if (args.has("-sim")) {
            run_sim(args);
        } else if (args.has("-batch")) {
            run_batch(args);
        } else if (args.has("-external")) {
            run_external(args);
        } else {
            std::cout << "missing program arguments, stopping execution (the evolutionary optimiser is "
                         "not part of this build; use -sim, -batch or -external)\n";
        }
        std::cout << "All done\n";
    } catch (const std::runtime_error& e) {
        std::cout << "runtime error caught: " << e.what() << "\n";
        return 1;
    } catch (const std::exception& e) {
        std::cout << "std exception caught: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
