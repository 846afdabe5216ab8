// ekg_model.cpp -- host FP64 parts of one evaluation that sit around the GPU
// simulation: the parameter vector -> per-layer WohlfartPlus coefficients
// (sim.cpp:751-916, with the transmural interpolation of sim.cpp:91-313 and
// nonlinearFit.h:92-168) and the comparison of a simulated ECG with a target
// (vectorMath.h:179-404, sim.cpp:640-660) and the target resampling
// (vectorMath.h:209-227).
//
// These run once per evaluation on a handful of numbers; they are plain C++
// and round like the reference build (no contraction, glibc exp/log/pow), so
// the iterative interpolation reaches the same coefficients bit for bit.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <set>
#include <vector>

#include "ekg_engine.hpp"
#include "ekg_model.hpp"

namespace ekg {

double ap_value(const double* k, double t)
{
    // WohlfartPlus::operator[] (Wohlfart.h:195-203); the operation order is
    // kept because callers compare FP64 results bit for bit.
    const double act = 1.0 / (1.0 + std::exp(-k[1] * t));
    const double plateau = k[2] * ((1.0 - k[3]) * std::exp(-k[4] * t) + k[3]);
    const double shift = std::log(std::pow(2.0, (k[7] / k[6])) - 1);
    const double repol = 1 - std::pow((1 + std::exp(-k[7] * (t - k[8]) + shift)), -(k[6] / k[7]));
    return act * plateau * (std::exp(-k[5] * t) * repol) + k[0];
}

double ap_apd90(const double* k)
{
    // WohlfartPlus::apd90 (Wohlfart.h:206-223): the last downward crossing of
    // k0 + 0.1 k2 on the 1 ms grid 0..999, placed by the reference's weighting
    // (a*wa + b*wb) / (wa + wb) with wa = d[a] - target, wb = target - d[b].
    const double target = k[0] + k[2] * 0.1;
    double prev = ap_value(k, 0.0);
    double when = -1.0;
    for (int i = 1; i < 1000; ++i) {
        const double cur = ap_value(k, (double)i);
        if (prev > target && cur <= target) {
            const double wa = prev - target, wb = target - cur;
            when = (double(i - 1) * wa + double(i) * wb) / (wa + wb);
        }
        prev = cur;
    }
    return when;
}

namespace {

inline double sq(double x) { return x * x; }

struct Ap {
    double k[9];
    double operator()(double t) const { return ap_value(k, t); }  // activation time 0
};

// size_t of a double the way the reference's x86-64 build converts a
// negative value (through a signed 64-bit integer), e.g. ceil(-1) -> SIZE_MAX.
inline size_t to_index(double v) { return v < 0 ? (size_t)(long long)v : (size_t)v; }

// Shape matcher between two border APs (WohlfartInterpolationEvaluator,
// sim.cpp:91-313): 15 line segments join points of equal relative arc length
// on the two APs (from 10 ms to past APD90); for a ratio r the matched AP
// should pass through the point at fraction r along each segment.
class ShapeMatcher {
public:
    void set_borders(const Ap& a, const Ap& b)
    {
        static const size_t kPoints = 15, kStart = 10, kLen = 700;
        static const double kXScale = 0.1;
        seg_.clear();
        const double apd_a = ap_apd90(a.k), apd_b = ap_apd90(b.k);
        // Samples 0..700; the reference keeps 0..699 and, for an AP whose
        // arc-length targets are not met before 699 ms, reads one sample past
        // its buffer (sim.cpp:219-226); here that sample is the AP at 700 ms.
        std::vector<double> ya(kLen + 1), yb(kLen + 1);
        for (size_t i = 0; i <= kLen; ++i) {
            ya[i] = a((double)i);
            yb[i] = b((double)i);
        }
        double len_a = 0.0, len_b = 0.0;
        for (size_t i = kStart + 1; i < std::min(kLen, to_index(std::ceil(apd_a))); ++i)
            len_a += std::sqrt(sq(ya[i] - ya[i - 1]) + kXScale);
        for (size_t i = kStart; i < std::min(kLen, to_index(std::ceil(apd_b))); ++i)
            len_b += std::sqrt(sq(yb[i] - yb[i - 1]) + kXScale);

        double reached_a = 0.0, reached_b = 0.0;
        size_t ia = kStart, ib = kStart;
        for (size_t p = 0; p < kPoints; ++p) {
            double goal = p * len_a / (kPoints - 2);
            for (double l = reached_a; (l < goal) && (ia < kLen); ++ia)
                l += std::sqrt(sq(ya[ia + 1] - ya[ia]) + kXScale);
            reached_a = goal;
            goal = p * len_b / (kPoints - 2);
            for (double l = reached_b; (l < goal) && (ib < kLen); ++ib)
                l += std::sqrt(sq(yb[ib + 1] - yb[ib]) + kXScale);
            reached_b = goal;

            Seg s;
            s.x1 = (double)ia;
            s.x2 = (double)ib;
            if (ia == ib) s.x2 += 0.001;
            s.k = (ya[ia] - yb[ib]) / (s.x1 - s.x2);
            s.n = yb[ib] - s.k * (double)ib;
            seg_.push_back(s);
        }
    }

    void set_ratio(double r)
    {
        pts_.clear();
        for (const Seg& s : seg_) {
            const double x = s.x1 + (s.x2 - s.x1) * r;
            pts_.push_back(x);
            pts_.push_back(s.k * x + s.n);
        }
    }

    // Sum of squared misses over the matched points.
    double operator()(const Ap& ap) const
    {
        double sum = 0;
        for (size_t i = 0; i < pts_.size(); i += 2) sum += sq(ap(pts_[i]) - pts_[i + 1]);
        return sum;
    }

private:
    struct Seg {
        double k, n, x1, x2;
    };
    std::vector<Seg> seg_;
    std::vector<double> pts_;
};

// Coordinate-wise steepest descent with per-coordinate adaptive steps
// (nonlinearFit.h:92-168): forward-difference gradient with offsets d/1000,
// a step of `step * move[i]` against the gradient sign, moves halved on a
// gradient sign flip and grown by 1.5 while the gradient stays large, the
// global step halved after a rejected move without a flip.
void descend(const ShapeMatcher& f, Ap& x0, const Ap& d, double step, double eps, int iterations)
{
    Ap grad = x0, old_grad = x0, move = d;
    for (double& g : old_grad.k) g = 0;
    double y0 = f(x0);
    for (; (iterations > 0) && (y0 > eps); --iterations) {
        bool flipped = false;
        for (int i = 0; i < 9; ++i) {
            if (d.k[i] != 0) {
                Ap x1 = x0;
                x1.k[i] += d.k[i] * .001;
                grad.k[i] = (f(x1) - y0) / (d.k[i] * .001);
                if (grad.k[i] * old_grad.k[i] < 0) {
                    move.k[i] *= 0.5;
                    flipped = true;
                } else if (std::fabs(grad.k[i]) > 0.75 * std::fabs(old_grad.k[i])) {
                    move.k[i] *= 1.5;
                }
            } else {
                grad.k[i] = 0;
            }
        }
        old_grad = grad;
        Ap x1 = x0;
        for (int i = 0; i < 9; ++i) x1.k[i] -= step * ((grad.k[i] > 0) ? move.k[i] : -move.k[i]);
        const double y1 = f(x1);
        if (y1 < y0) {
            y0 = y1;
            x0 = x1;
        } else if (!flipped) {
            step *= 0.5;
        }
    }
}

// Descent offsets for the interpolated layers (sim.cpp:798, 880).
const double kDescentOffsets[9] = {0, 0, 0, 0.001, 0, 0.00005, 0.0005, 0.01, 0.2};

// k = a (1 - r) + b r for all nine coefficients (sim.cpp:705-709).
void blend(Ap& out, const Ap& a, const Ap& b, double r)
{
    for (int i = 0; i < 9; ++i) out.k[i] = a.k[i] * (1 - r) + b.k[i] * r;
}

}  // namespace

int ap_num_genes(const ApModel& m)
{
    // std::set<char> of the free indices (sim.cpp:492-509): duplicates count
    // once, every distinct entry counts even if it names no coefficient.
    std::set<char> s;
    for (int i = 0; i < m.n_free; ++i) s.insert((char)m.free_k[i]);
    return (int)s.size() * (m.interpolation == EKG_INTERP_ENDO_EPI ? 2 : 3);
}

void layer_aps(const ApModel& m, int n_layers, const double* genes, int n_genes, double* layer_k,
               double* violation)
{
    if (!m.base_k || m.n_base < 1) fail(EKG_EINVAL, "need at least one base AP");
    if (!m.k_min || !m.k_max) fail(EKG_EINVAL, "k min / k max missing");
    if (m.n_free < 0 || (m.n_free > 0 && !m.free_k)) fail(EKG_EINVAL, "bad free-k list");
    if (m.interpolation != EKG_INTERP_ENDO_EPI && m.interpolation != EKG_INTERP_ENDO_MID_EPI)
        fail(EKG_EINVAL, "unknown interpolation type");
    if (n_layers < 1) fail(EKG_EINVAL, "need at least one layer");
    if (!layer_k) fail(EKG_EINVAL, "null output");
    std::set<char> free_set;
    for (int i = 0; i < m.n_free; ++i) free_set.insert((char)m.free_k[i]);
    const int used = ap_num_genes(m);
    if (n_genes < used || (used > 0 && !genes))
        fail(EKG_EINVAL, "parameter vector shorter than the free Wohlfart coefficients need");

    int gi = 0;
    double viol = 0.0;
    // A border AP: base coefficients with the free ones taken from the
    // vector in coefficient order; each free value outside [k min, k max]
    // adds its distance to the violation (sim.cpp:582-596).
    auto border = [&](const double* base) {
        Ap ap;
        std::memcpy(ap.k, base, sizeof ap.k);
        for (int i = 0; i < 9; ++i)
            if (free_set.count((char)i)) {
                ap.k[i] = genes[gi++];
                if (ap.k[i] < m.k_min[i])
                    viol += m.k_min[i] - ap.k[i];
                else if (ap.k[i] > m.k_max[i])
                    viol += ap.k[i] - m.k_max[i];
            }
        return ap;
    };
    const double* first_base = m.base_k;
    const double* last_base = m.base_k + 9 * (m.n_base - 1);
    const size_t n = (size_t)n_layers;
    std::vector<Ap> aps(n);
    Ap d;
    std::memcpy(d.k, kDescentOffsets, sizeof d.k);
    ShapeMatcher match;

    if (m.interpolation == EKG_INTERP_ENDO_EPI) {
        // sim.cpp:751-821: first and last layers from the vector, the others
        // matched at ratio i/(n-1) between them.
        aps.front() = border(first_base);
        aps.back() = border(last_base);
        match.set_borders(aps.front(), aps.back());
        for (size_t i = 1; i + 1 < n; ++i) {
            const double r = i / double(n - 1);
            match.set_ratio(r);
            aps[i] = aps[i - 1];
            blend(aps[i], aps.front(), aps.back(), r);
            descend(match, aps[i], d, 0.5, 1e-3, 100);
        }
    } else {
        // sim.cpp:825-916: endo, mid (at floor(pos (n-1) + 0.5)) and epi from
        // the vector, both mid and epi starting from the last base AP; the
        // layers between are matched on each side of the mid layer.
        const size_t mid = (size_t)std::floor(m.mid_position * (n - 1) + 0.5);
        if (mid >= n) fail(EKG_EINVAL, "mid AP position outside [0, 1]");
        aps.front() = border(first_base);
        aps[mid] = border(last_base);
        aps.back() = border(last_base);
        for (size_t i = 1; i + 1 < n; ++i) {
            if (i < mid) {
                const double r = i / double(mid);
                match.set_borders(aps.front(), aps[mid]);
                match.set_ratio(r);
                aps[i] = aps[i - 1];
                blend(aps[i], aps.front(), aps[mid], r);
            } else if (i > mid) {
                const double r = (i - mid) / double(n - mid - 1);
                if (i - mid == 1) match.set_borders(aps[mid], aps.back());
                match.set_ratio(r);
                aps[i] = aps[i - 1];
                blend(aps[i], aps[mid], aps.back(), r);
            }
            if (i != mid) descend(match, aps[i], d, 0.5, 1e-3, 100);
        }
    }
    for (size_t l = 0; l < n; ++l) std::memcpy(layer_k + 9 * l, aps[l].k, sizeof aps[l].k);
    if (violation) *violation = viol;
}

namespace {

// Sum, mean/variance and covariance in the reference's accumulation order
// (vectorMath.h:145-186, 228-237).
double range_sum(const double* v, size_t a, size_t b)
{
    double s = 0;
    for (size_t i = a; i < b; ++i) s += v[i];
    return s;
}

void mean_var(double& mean, double& var, const double* v, size_t a, size_t b)
{
    const double inv = 1.0 / double(b - a);
    mean = range_sum(v, a, b) * inv;
    var = 0;
    for (size_t i = a; i < b; ++i) var += sq(v[i] - mean);
    var *= inv;
}

}  // namespace

double criterion(int mode, const double* a, size_t na, const double* b, size_t nb, double b_offset)
{
    // All comparisons use offset 0 (sim.cpp:640-660): the first
    // min(na, nb) samples of both curves.
    const size_t len = std::min(na, nb);
    switch (mode) {
    case EKG_CMP_CORRELATION: {
        // 1 - Pearson r (statisticalCorrelationCoeff, vectorMath.h:277-307)
        double ma, va, mb, vb;
        mean_var(ma, va, a, 0, len);
        mean_var(mb, vb, b, 0, len);
        double cov = 0;
        for (size_t i = 0; i < len; ++i) cov += (a[i] - ma) * (b[i] - mb);
        cov /= double(len);
        return 1.0 - cov / (std::sqrt(va) * std::sqrt(vb));
    }
    case EKG_CMP_VECTOR_CORRELATION: {
        // 1 - a.b / (|a| |b|) (vectorCorrelationCoeff, vectorMath.h:338-357)
        double ab = 0, aa = 0, bb = 0;
        for (size_t i = 0; i < len; ++i) ab += a[i] * b[i];
        for (size_t i = 0; i < len; ++i) aa += a[i] * a[i];
        for (size_t i = 0; i < len; ++i) bb += b[i] * b[i];
        return 1.0 - ab / (std::sqrt(aa) * std::sqrt(bb));
    }
    case EKG_CMP_RMS: {
        // sqrt(mean((a - b)^2)) (rms, vectorMath.h:311-333)
        std::vector<double> d2(len);
        for (size_t i = 0; i < len; ++i) {
            const double d = a[i] - b[i];
            d2[i] = d * d;
        }
        double mean, var;
        mean_var(mean, var, d2.data(), 0, len);
        return std::sqrt(mean);
    }
    case EKG_CMP_DEV_FROM_LINEAR: {
        // standard deviation of (a/range(a) + off) / (b + off), with the range
        // over all of a (devFromLinear, vectorMath.h:363-401)
        double lo = 0, hi = 0;
        if (na > 0) {
            lo = hi = a[0];
            for (size_t i = 1; i < na; ++i) {
                if (a[i] < lo)
                    lo = a[i];
                else if (a[i] > hi)
                    hi = a[i];
            }
        }
        const double scale = 1.0 / (hi - lo);
        std::vector<double> q(len);
        for (size_t i = 0; i < len; ++i) q[i] = (a[i] * scale + b_offset) / (b[i] + b_offset);
        double mean, var;
        mean_var(mean, var, q.data(), 0, len);
        return var == 0 ? 1 / 1e-30 : std::sqrt(var);
    }
    default:
        fail(EKG_EINVAL, "unknown comparison mode");
    }
    return 0;
}

size_t resample(const double* src, size_t n, double factor, int start, double* dest, size_t cap)
{
    // vectorMath.h:209-227: linear interpolation at i / factor + start; a
    // factor of exactly 1 copies the source and ignores start.  The length
    // arithmetic is unsigned, as in the reference.
    if (n == 0) fail(EKG_EINVAL, "empty source");
    if (factor == 1) {
        for (size_t i = 0; i < n && i < cap; ++i) dest[i] = src[i];
        return n;
    }
    const size_t span = n - 1 - (size_t)start;
    const double len = std::ceil((double)span * factor);
    if (!(len >= 0) || len > 1e9) fail(EKG_EINVAL, "resampled length out of range");
    const size_t out = 1u + (size_t)len;
    for (size_t i = 0; i < out && i < cap; ++i) {
        const double pos = i / factor + start;
        const size_t f = to_index(std::floor(pos));
        const size_t c = (size_t)(1.0 + f);
        dest[i] = c < n ? src[c] * (pos - f) + src[f] * (c - pos) : src[n - 1];
    }
    return out;
}

}  // namespace ekg
