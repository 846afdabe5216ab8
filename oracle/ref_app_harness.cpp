// ref_app_harness.cpp -- C entry points around the reference application's
// host-side model code: the transmural AP interpolation (sim.cpp:72-313 with
// nonlinearFit.h:92-168), WohlfartPlus::apd90 (Wohlfart.h:206-223), the ECG
// comparisons (vectorMath.h:179-404) and target resampling.  Compiled from the
// reference's own sources (oracle/Makefile, target `ref`); test
// infrastructure only, used to pin the product's ekg_layer_aps and
// ekg_criterion bit for bit.  Nothing in the product links or loads it.
#include "sim.cpp"

#include <cstring>

namespace {

SimLib::ActionPotential make_ap(const double* k)
{
    SimLib::ActionPotential ap;
    ap.init(k, 0);
    return ap;
}

// combineAps (sim.cpp:705-709) is a member of the reference's hidden
// SimImplementation class; the same one-line blend is written out here.
void blend(SimLib::ActionPotential& d, const SimLib::ActionPotential& a, const SimLib::ActionPotential& b, double r)
{
    for (size_t i = 0; i < a.wohl.numParams; ++i) d.wohl.getK()[i] = a.wohl.getK()[i] * (1 - r) + b.wohl.getK()[i] * r;
}

}  // namespace

extern "C" {

// Layers of simUsingBorderAps (interp 1, sim.cpp:785-808) or
// simUsingBorderAndMidAps (interp 2, sim.cpp:867-905) from border APs that
// already hold the parameter-vector values.  out: [n_layers][9].
int ref_interp_layers(int interp, int n_layers, double mid_position, const double* front_k, const double* mid_k,
                      const double* back_k, double* out)
{
    const size_t n = (size_t)n_layers;
    std::vector<SimLib::ActionPotential> aps(n);
    double kd[] = {0, 0, 0, 0.001, 0, 0.00005, 0.0005, 0.01, 0.2};
    SimLib::ActionPotential d;
    d.init(kd, 0);
    WohlfartInterpolationEvaluator ev;
    if (interp == 1) {
        aps.front() = make_ap(front_k);
        aps.back() = make_ap(back_k);
        ev.setBorderAps(aps.front(), aps.back());
        for (size_t i = 1; i + 1 < n; ++i) {
            double ratio = i / double(n - 1);
            ev.setupRatio(ratio);
            aps[i] = aps[i - 1];
            blend(aps[i], aps.front(), aps.back(), ratio);
            steepestDescend(ev, aps[i], d, 0.5, 1e-3, 100);
        }
    } else {
        size_t mid = (size_t)floor(mid_position * (n - 1) + 0.5);
        aps.front() = make_ap(front_k);
        aps[mid] = make_ap(mid_k);
        aps.back() = make_ap(back_k);
        for (size_t i = 1; i + 1 < n; ++i) {
            if (i < mid) {
                double ratio = i / double(mid);
                ev.setBorderAps(aps.front(), aps[mid]);
                ev.setupRatio(ratio);
                aps[i] = aps[i - 1];
                blend(aps[i], aps.front(), aps[mid], ratio);
            } else if (i > mid) {
                double ratio = (i - mid) / double(n - mid - 1);
                if ((i - mid) == 1) ev.setBorderAps(aps[mid], aps.back());
                ev.setupRatio(ratio);
                aps[i] = aps[i - 1];
                blend(aps[i], aps[mid], aps.back(), ratio);
            }
            if (i != mid) steepestDescend(ev, aps[i], d, 0.5, 1e-3, 100);
        }
    }
    for (size_t l = 0; l < n; ++l) std::memcpy(out + 9 * l, aps[l].getK(), 9 * sizeof(double));
    return 0;
}

double ref_apd90(const double* k)
{
    WohlfartPlus w;
    w.setK(k);
    return w.apd90();
}

// Comparison value as SimImplementation::calculateFitness forms it
// (sim.cpp:640-660): 1 - r for the correlation modes, offset 0.
double ref_compare(int mode, const double* a, int64_t na, const double* b, int64_t nb, double bofs)
{
    std::vector<double> va(a, a + na), vb(b, b + nb);
    ConvolutionResult<double> cr;
    switch (mode) {
    case 2:
        cr = statisticalCorrelationCoeff(va, vb, 0);
        cr.bestValue = 1.0 - cr.bestValue;
        break;
    case 4:
        cr = vectorCorrelationCoeff(va, vb, 0);
        cr.bestValue = 1.0 - cr.bestValue;
        break;
    case 1:
        cr = rms(va, vb, 0);
        break;
    default:
        cr = devFromLinear(va, vb, bofs, 0);
        break;
    }
    return cr.bestValue;
}

// resample (vectorMath.h:209-227); returns the output length (written up to cap).
int64_t ref_resample(const double* src, int64_t n, double factor, int start, double* dest, int64_t cap)
{
    std::vector<double> s(src, src + n), d;
    resample(s, d, factor, start);
    for (int64_t i = 0; i < (int64_t)d.size() && i < cap; ++i) dest[i] = d[i];
    return (int64_t)d.size();
}

}  // extern "C"
