// stone_width.cpp -- the girdle width that scales the metrics ray fan.
//
// measure_model_width (render_setup/stone_width.rs:27-42) reduces a plane
// arrangement to the short axis of its vertices' x/z bounding box
// (geometry/stone_metrics/measure.rs:79-150).  The vertices are every
// well-conditioned triple of the (deduplicated, power-of-two normalised)
// planes plus the six faces of a 64-unit blank whose solved point satisfies
// every half-space within 1e-5; a feasible point on the blank means the
// planes do not close a finite solid.  All of it is f64 host work done once
// per design (the fan itself is traced on the device).
//
// Pairs whose common line misses the solid by more than the solve's error
// bound are skipped before their inner loop (vertices.rs:141-191): a
// conservative prune that visits the same triples, so the bounding box is
// unchanged.  Designs are measured on a few host threads.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/ixg.h"

namespace {

struct D3 {
    double x, y, z;
};
inline D3 d3(double x, double y, double z) { return D3{x, y, z}; }
inline D3 operator-(D3 a, D3 b) { return d3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline D3 operator+(D3 a, D3 b) { return d3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline D3 operator*(D3 a, double s) { return d3(a.x * s, a.y * s, a.z * s); }
inline double dot(D3 a, D3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline D3 cross(D3 a, D3 b) { return d3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y); }
inline double absmax(D3 v) { return std::fmax(std::fmax(std::fabs(v.x), std::fabs(v.y)), std::fabs(v.z)); }

constexpr double kBlank = 64.0;        // meet_solver/mod.rs:218
constexpr double kEpsFeas = 1e-5;      // meet_solver/mod.rs:223
constexpr double kMinDet = 1e-6;       // meet_solver/mod.rs:235
constexpr double kDedup = 1e-6;        // stone_metrics/mod.rs:112
constexpr double kEpsFace = 2e-5;      // stone_metrics/mod.rs:109
constexpr double kPruneMinSin = 1e-3;  // vertices.rs:143
constexpr double kResidual = 1e-11;    // vertices.rs:151

// normals_coincide (cuts/asc_schedule.rs:335-339)
bool same_direction(D3 a, D3 b) {
    auto unit = [](D3 v) {
        double r = 1.0 / std::sqrt(dot(v, v));
        return (std::isfinite(r) && r > 0.0) ? v * r : d3(0, 0, 0);
    };
    return 1.0 - dot(unit(a), unit(b)) <= 2e-7;
}

// pow2_scale_norm (stone_metrics/mod.rs:152-174): the power of two nearest
// (in log terms) to a positive scale, 1 for anything else.
double pow2_scale(double s) {
    if (!(std::isfinite(s) && s > 0.0)) return 1.0;
    int e = 0;
    double m = std::frexp(s, &e);  // s = m * 2^e, m in [0.5, 1)
    // mantissa in [1, 2) is 2m, exponent e - 1
    int k = e - 1;
    if (2.0 * m >= 1.41421356237309504880) k += 1;
    k = std::min(std::max(k, -1022), 1023);
    return std::ldexp(1.0, k);
}

struct Plane {
    D3 n;
    double m;  // n . x <= m
};

// The triple solve (simd/triple_solve.rs): cofactor inverse times the offsets.
double solve3(const Plane& a, const Plane& b, const Plane& c, D3* v) {
    const D3 A = a.n, B = b.n, C = c.n;
    double t0x = B.y * C.z - B.z * C.y, t0y = C.y * A.z - C.z * A.y, t0z = A.y * B.z - A.z * B.y;
    double det = (A.x * t0x + B.x * t0y) + C.x * t0z;
    double t1x = B.z * C.x - B.x * C.z, t1y = C.z * A.x - C.x * A.z, t1z = A.z * B.x - A.x * B.z;
    double t2x = B.x * C.y - B.y * C.x, t2y = C.x * A.y - C.y * A.x, t2z = A.x * B.y - A.y * B.x;
    double inv = 1.0 / det;
    v->x = ((t0x * inv) * a.m + (t0y * inv) * b.m) + (t0z * inv) * c.m;
    v->y = ((t1x * inv) * a.m + (t1y * inv) * b.m) + (t1z * inv) * c.m;
    v->z = ((t2x * inv) * a.m + (t2y * inv) * b.m) + (t2z * inv) * c.m;
    return det;
}

// True only if no triple through planes a and b can be feasible: the line
// where they meet, clipped by every other half-space widened by the
// feasibility slack plus the solve's residual bound, is empty.
bool pair_misses(const std::vector<Plane>& all, size_t a, size_t b) {
    const D3 dir = cross(all[a].n, all[b].n);
    const double sin = std::sqrt(dot(dir, dir));
    if (sin < kPruneMinSin) return false;
    const D3 origin = (cross(all[b].n, dir) * all[a].m + cross(dir, all[a].n) * all[b].m) * (1.0 / (sin * sin));
    const D3 unit = dir * (1.0 / sin);
    const double margin = kEpsFeas + 2.0 * kResidual / kMinDet / sin;
    double lo = -INFINITY, hi = INFINITY;
    for (size_t c = 0; c < all.size(); ++c) {
        if (c == a || c == b) continue;
        const double slope = dot(all[c].n, unit);
        const double room = all[c].m + margin - dot(all[c].n, origin);
        if (slope > 0.0) hi = std::min(hi, room / slope);
        else if (slope < 0.0) lo = std::max(lo, room / slope);
        else if (room < 0.0) return true;
        if (lo > hi) return true;
    }
    return false;
}

double face_area(const Plane& p, const std::vector<D3>& verts) {
    std::vector<D3> on;
    for (const D3& v : verts)
        if (std::fabs(dot(p.n, v) - p.m) <= kEpsFace) on.push_back(v);
    if (on.size() < 3) return 0.0;
    const D3 n = p.n;
    D3 seed = (std::fabs(n.x) <= std::fabs(n.y) && std::fabs(n.x) <= std::fabs(n.z)) ? d3(1, 0, 0)
              : (std::fabs(n.y) <= std::fabs(n.z))                                ? d3(0, 1, 0)
                                                                                  : d3(0, 0, 1);
    D3 bu = seed - n * dot(n, seed);
    bu = bu * (1.0 / std::sqrt(dot(bu, bu)));
    const D3 bv = cross(n, bu);
    D3 c = d3(0, 0, 0);
    for (const D3& v : on) c = c + v;
    c = c * (1.0 / (double)on.size());
    std::vector<std::pair<double, D3>> ring;
    for (const D3& v : on) ring.push_back({std::atan2(dot(bv, v - c), dot(bu, v - c)), v});
    std::sort(ring.begin(), ring.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
    D3 s = d3(0, 0, 0);
    for (size_t i = 0; i < ring.size(); ++i)
        s = s + cross(ring[i].second - c, ring[(i + 1) % ring.size()].second - c);
    return 0.5 * std::fabs(dot(n, s));
}

// measure_solid's width_axis; false when the planes do not measure.
bool measure_width(const ixg_facet_plane* planes, uint32_t n, double* width) {
    if (n == 0) return false;
    std::vector<Plane> all;
    all.reserve(n + 6);
    for (uint32_t i = 0; i < n; ++i) {  // dedup_planes (vertices.rs:67-79)
        Plane p{d3(planes[i].normal[0], planes[i].normal[1], planes[i].normal[2]), -(double)planes[i].d};
        bool dup = false;
        for (const Plane& q : all)
            if (same_direction(p.n, q.n) && std::fabs(p.m - q.m) < 1e-9) {
                dup = true;
                break;
            }
        if (!dup) all.push_back(p);
    }
    double amax = 0.0;  // max_abs_offset, NaN-propagating
    for (const Plane& p : all) amax = (std::isnan(amax) || std::isnan(p.m)) ? NAN : std::fmax(amax, std::fabs(p.m));
    const double scale = pow2_scale(amax);
    for (Plane& p : all) p.m /= scale;
    const size_t real = all.size();
    const D3 axes[6] = {d3(1, 0, 0), d3(-1, 0, 0), d3(0, 1, 0), d3(0, -1, 0), d3(0, 0, 1), d3(0, 0, -1)};
    for (const D3& ax : axes) all.push_back(Plane{ax, kBlank});
    const size_t p = all.size();

    std::vector<D3> verts;
    for (size_t a = 0; a < p; ++a)
        for (size_t b = a + 1; b < p; ++b) {
            if (pair_misses(all, a, b)) continue;
            for (size_t c = b + 1; c < p; ++c) {
                D3 v;
                const double det = std::fabs(solve3(all[a], all[b], all[c], &v));
                if (std::isnan(det) || det < kMinDet) continue;
                if (!(std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z))) continue;
                const double am = absmax(v);
                if (am > kBlank + 1.0) continue;
                bool viol = false;
                for (size_t q = 0; q < p && !viol; ++q) viol = dot(all[q].n, v) - all[q].m > kEpsFeas;
                if (viol) continue;
                if (am > kBlank - 1.0) return false;  // reaches the blank: unbounded
                bool dup = false;
                for (const D3& w : verts)
                    if (absmax(w - v) < kDedup) {
                        dup = true;
                        break;
                    }
                if (!dup) verts.push_back(v);
            }
        }
    if (verts.size() < 4) return false;
    double vol = 0.0;
    for (size_t i = 0; i < real; ++i) vol += all[i].m * face_area(all[i], verts) / 3.0;
    if (!(std::isfinite(vol) && vol > 0.0)) return false;
    double x0 = INFINITY, x1 = -INFINITY, z0 = INFINITY, z1 = -INFINITY;
    for (const D3& v : verts) {
        x0 = std::fmin(x0, v.x);
        x1 = std::fmax(x1, v.x);
        z0 = std::fmin(z0, v.z);
        z1 = std::fmax(z1, v.z);
    }
    const double dx = x1 - x0, dz = z1 - z0;
    *width = (dx <= dz ? dx : dz) * scale;
    return true;
}

}  // namespace

extern "C" ixg_status ixg_model_width(const ixg_facet_plane* planes, uint32_t n_planes, double* width) {
    if ((!planes && n_planes) || !width) return IXG_INVALID;
    return measure_width(planes, n_planes, width) ? IXG_OK : IXG_DECLINED;
}

// girdle_half_width (color/metrics/fan.rs:49-60) for each design of a plane
// table: width / 2 as f32, or 1.0 when the design does not measure.
extern "C" void ixg_internal_half_widths(const ixg_facet_plane* planes, const ixg_design* designs, uint32_t n,
                                         float* out) {
    auto one = [&](uint32_t d) {
        double w = 0.0;
        bool ok = measure_width(planes + designs[d].plane_offset, designs[d].n_planes, &w);
        out[d] = (ok && std::isfinite(w) && w > 0.0) ? (float)(w * 0.5) : 1.0f;
    };
    unsigned hw = std::thread::hardware_concurrency();
    unsigned threads = std::min<unsigned>(std::max(1u, std::min(hw, 16u)), n);
    if (threads <= 1) {
        for (uint32_t d = 0; d < n; ++d) one(d);
        return;
    }
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < threads; ++t)
        pool.emplace_back([&, t] {
            for (uint32_t d = t; d < n; d += threads) one(d);
        });
    for (auto& th : pool) th.join();
}
