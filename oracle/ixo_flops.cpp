// ixo_flops.cpp -- the FLOP-counting build of the oracle (test and
// measurement infrastructure only): ixo.c compiled as C++ with every `float`
// a counting f32 type, so a film or metric call returns, besides its result,
// the exact number of floating-point operations the reference's algorithm
// (as the oracle restates it, operation for operation) performs on those
// paths.  The values stay f32 and bit-identical to the FP32 oracle's, so the
// paths and their decisions are the FP32 oracle's own.
//
// Counting rule: add, subtract, multiply, divide, square root and every
// transcendental count 1, a fused multiply-add counts 2; negation, absolute
// value, min / max / clamp, comparisons, conversions and integer work count 0.
// The calls are single-threaded (the counter is one process-wide tally).
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <type_traits>

#include "../include/ixg.h"
#include "ixo.h"

namespace {
uint64_t g_flops = 0;
}
#define IXO_SCALAR_ONLY 1
#define IXO_NO_PROBE 1
#define IXO_FLOPS(n) (g_flops += (uint64_t)(n))

struct CF {
    float v;
    CF() = default;
    constexpr CF(float x) : v(x) {}
    CF(double x) : v((float)x) {}
    CF(int x) : v((float)x) {}
    CF(unsigned x) : v((float)x) {}
    CF(long x) : v((float)x) {}
    CF(unsigned long x) : v((float)x) {}
    CF(long long x) : v((float)x) {}
    CF(unsigned long long x) : v((float)x) {}
    explicit operator float() const { return v; }
    explicit operator double() const { return v; }
    explicit operator int() const { return (int)v; }
    explicit operator unsigned() const { return (unsigned)v; }
    explicit operator long() const { return (long)v; }
    explicit operator unsigned long() const { return (unsigned long)v; }
    explicit operator long long() const { return (long long)v; }
    explicit operator unsigned long long() const { return (unsigned long long)v; }
    explicit operator bool() const { return v != 0.0f; }
    CF& operator+=(CF o) { ++g_flops; v += o.v; return *this; }
    CF& operator-=(CF o) { ++g_flops; v -= o.v; return *this; }
    CF& operator*=(CF o) { ++g_flops; v *= o.v; return *this; }
    CF& operator/=(CF o) { ++g_flops; v /= o.v; return *this; }
};
static_assert(sizeof(CF) == 4 && std::is_trivial<CF>::value, "CF is an f32 in storage");

template <class T>
using arith = typename std::enable_if<std::is_arithmetic<T>::value, int>::type;
inline CF cfv(CF a) { return a; }
template <class T, arith<T> = 0>
inline CF cfv(T a) { return CF((float)a); }

#define IXO_CF_BINOP(OP)                                                                 \
    inline CF operator OP(CF a, CF b) { ++g_flops; return CF(a.v OP b.v); }             \
    template <class T, arith<T> = 0>                                                     \
    inline CF operator OP(CF a, T b) { ++g_flops; return CF(a.v OP (float)b); }          \
    template <class T, arith<T> = 0>                                                     \
    inline CF operator OP(T a, CF b) { ++g_flops; return CF((float)a OP b.v); }
IXO_CF_BINOP(+)
IXO_CF_BINOP(-)
IXO_CF_BINOP(*)
IXO_CF_BINOP(/)
#define IXO_CF_CMP(OP)                                                                   \
    inline bool operator OP(CF a, CF b) { return a.v OP b.v; }                           \
    template <class T, arith<T> = 0>                                                     \
    inline bool operator OP(CF a, T b) { return a.v OP (float)b; }                       \
    template <class T, arith<T> = 0>                                                     \
    inline bool operator OP(T a, CF b) { return (float)a OP b.v; }
IXO_CF_CMP(<)
IXO_CF_CMP(>)
IXO_CF_CMP(<=)
IXO_CF_CMP(>=)
IXO_CF_CMP(==)
IXO_CF_CMP(!=)
inline CF operator-(CF a) { return CF(-a.v); }
inline CF operator+(CF a) { return a; }

// the f32 libm calls of ixo.c
template <class A, class B, class C>
inline CF cf_fma(A a, B b, C c) { g_flops += 2; return CF(::fmaf(cfv(a).v, cfv(b).v, cfv(c).v)); }
#define IXO_CF_UNARY(NAME, FN, COST) \
    template <class A>               \
    inline CF NAME(A a) { g_flops += (COST); return CF(FN(cfv(a).v)); }
IXO_CF_UNARY(cf_sqrt, ::sqrtf, 1)
IXO_CF_UNARY(cf_cos, ::cosf, 1)
IXO_CF_UNARY(cf_sin, ::sinf, 1)
IXO_CF_UNARY(cf_tan, ::tanf, 1)
IXO_CF_UNARY(cf_acos, ::acosf, 1)
IXO_CF_UNARY(cf_atan, ::atanf, 1)
IXO_CF_UNARY(cf_exp, ::expf, 1)
IXO_CF_UNARY(cf_log, ::logf, 1)
IXO_CF_UNARY(cf_floor, ::floorf, 1)
IXO_CF_UNARY(cf_fabs, ::fabsf, 0)
template <class A, class B>
inline CF cf_atan2(A a, B b) { ++g_flops; return CF(::atan2f(cfv(a).v, cfv(b).v)); }
template <class A, class B>
inline CF cf_pow(A a, B b) { ++g_flops; return CF(::powf(cfv(a).v, cfv(b).v)); }
template <class A, class B>
inline CF cf_fmod(A a, B b) { ++g_flops; return CF(::fmodf(cfv(a).v, cfv(b).v)); }
template <class A, class B>
inline CF cf_fmax(A a, B b) { return CF(::fmaxf(cfv(a).v, cfv(b).v)); }
template <class A, class B>
inline CF cf_fmin(A a, B b) { return CF(::fminf(cfv(a).v, cfv(b).v)); }
template <class A, class B>
inline CF cf_copysign(A a, B b) { return CF(::copysignf(cfv(a).v, cfv(b).v)); }
inline bool isfinite(CF a) { return std::isfinite(a.v); }
inline bool isnan(CF a) { return std::isnan(a.v); }
inline bool isinf(CF a) { return std::isinf(a.v); }
inline bool signbit(CF a) { return std::signbit(a.v); }

#define float CF
#define fmaf cf_fma
#define sqrtf cf_sqrt
#define fabsf cf_fabs
#define cosf cf_cos
#define sinf cf_sin
#define tanf cf_tan
#define acosf cf_acos
#define atanf cf_atan
#define atan2f cf_atan2
#define floorf cf_floor
#define expf cf_exp
#define logf cf_log
#define powf cf_pow
#define fmodf cf_fmod
#define fmaxf cf_fmax
#define fminf cf_fmin
#define copysignf cf_copysign

#define ixo_accumulate ixof_accumulate
#define ixo_asc_planes ixof_asc_planes
#define ixo_band_evaluate ixof_band_evaluate
#define ixo_blackbody ixof_blackbody
#define ixo_camera_new ixof_camera_new
#define ixo_cie_cmf ixof_cie_cmf
#define ixo_classify_ray ixof_classify_ray
#define ixo_cosine_hemisphere ixof_cosine_hemisphere
#define ixo_cut_round_brilliant ixof_cut_round_brilliant
#define ixo_cut_standard_round_brilliant ixof_cut_standard_round_brilliant
#define ixo_d65 ixof_d65
#define ixo_dispersion_evaluate ixof_dispersion_evaluate
#define ixo_exp_lane ixof_exp_lane
#define ixo_exp_lane8 ixof_exp_lane8
#define ixo_generate_ray ixof_generate_ray
#define ixo_girdle_half_width ixof_girdle_half_width
#define ixo_hash_u32 ixof_hash_u32
#define ixo_hg_direction ixof_hg_direction
#define ixo_image_offset ixof_image_offset
#define ixo_image_scintillation_pct ixof_image_scintillation_pct
#define ixo_intersect ixof_intersect
#define ixo_low_discrepancy_base2 ixof_low_discrepancy_base2
#define ixo_metric_image ixof_metric_image
#define ixo_model_width ixof_model_width
#define ixo_pose_metrics ixof_pose_metrics
#define ixo_radical_inverse_base ixof_radical_inverse_base
#define ixo_rgb_to_spectral ixof_rgb_to_spectral
#define ixo_sample_draws ixof_sample_draws
#define ixo_sample_studio_environment ixof_sample_studio_environment
#define ixo_scintillation_index ixof_scintillation_index
#define ixo_scintillation_index_at ixof_scintillation_index_at
#define ixo_scintillation_indices ixof_scintillation_indices
#define ixo_scintillation_spatial ixof_scintillation_spatial
#define ixo_scintillation_spatial_at ixof_scintillation_spatial_at
#define ixo_slab_scan ixof_slab_scan
#define ixo_shading_normal_near_edge ixof_shading_normal_near_edge
#define ixo_solid_outline ixof_solid_outline
#define ixo_temporal_cameras ixof_temporal_cameras
#define ixo_trace_paths ixof_trace_paths
#define ixo_trace_ray ixof_trace_ray
#define ixo_trace_sample ixof_trace_sample
#define ixo_uniaxial_energy ixof_uniaxial_energy
#define ixo_uniaxial_entry_r ixof_uniaxial_entry_r
#define ixo_white_balance ixof_white_balance
#define ixo_wrapped_hero_wavelengths ixof_wrapped_hero_wavelengths

#include "ixo.c"

#undef float

extern "C" {

// ixo_accumulate on one thread; *flops = the operations it performed.
void ixo_flops_accumulate(const ixg_scene* scene, uint32_t sample_offset, uint32_t spp, double* accum,
                          uint64_t* flops) {
    g_flops = 0;
    ixof_accumulate(scene, sample_offset, spp, accum, 1);
    *flops = g_flops;
}

// ixo_metric_image; *flops = the operations it performed.
int ixo_flops_metric_image(const ixg_metric_request* r, ixg_optical_pixel* pixels, ixg_metric_totals* tot,
                           uint64_t* flops) {
    g_flops = 0;
    const int st = ixof_metric_image(r, pixels, tot);
    *flops = g_flops;
    return st;
}

// ixo_intersect alone (the slab scan), for the counting rule's own check.
int32_t ixo_flops_intersect(const ixg_facet_plane* planes, uint32_t n, const float origin[3], const float dir[3],
                            float* t_out, uint64_t* flops) {
    g_flops = 0;
    CF t;
    const int32_t f = ixof_intersect(planes, n, reinterpret_cast<const CF*>(origin),
                                     reinterpret_cast<const CF*>(dir), &t);
    *t_out = t.v;
    *flops = g_flops;
    return f;
}

}  // extern "C"
