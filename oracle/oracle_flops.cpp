// oracle_flops.cpp -- TEST INFRASTRUCTURE ONLY (measurement).
//
// Counts the floating-point arithmetic of the oracle's query::contact
// (parry_oracle_contact.hpp) pair by pair: the restatement is a template over
// its scalar, and here it is instantiated with CountedF, an f32 that counts
// every add/sub, mul, div and sqrt it performs (and, separately, comparisons
// and min/max/abs, which are VALU work but not flops).  bench.py divides the
// flop count by the GPU's contact time to give the FP32 fraction of the
// vector peak.  The values computed are exactly the f32 restatement's.
#include <cmath>
#include <cstdint>
#include <limits>
#include <type_traits>

namespace oracle {

struct FlopCount {
    uint64_t add = 0, mul = 0, div = 0, sqrt = 0, cmp = 0, other = 0;
};
inline thread_local FlopCount g_flops;

struct CountedF {
    float v;
    CountedF() = default;
    template <class U, class = std::enable_if_t<std::is_arithmetic<U>::value>>
    constexpr CountedF(U x) : v((float)x) {}
    explicit operator float() const { return v; }
    explicit operator double() const { return v; }
    CountedF operator-() const { return CountedF(-v); }  // a sign flip, folded into the consumer
    CountedF &operator+=(CountedF b) { ++g_flops.add; v += b.v; return *this; }
    CountedF &operator-=(CountedF b) { ++g_flops.add; v -= b.v; return *this; }
    CountedF &operator*=(CountedF b) { ++g_flops.mul; v *= b.v; return *this; }
    CountedF &operator/=(CountedF b) { ++g_flops.div; v /= b.v; return *this; }
};
inline CountedF operator+(CountedF a, CountedF b) { ++g_flops.add; return CountedF(a.v + b.v); }
inline CountedF operator-(CountedF a, CountedF b) { ++g_flops.add; return CountedF(a.v - b.v); }
inline CountedF operator*(CountedF a, CountedF b) { ++g_flops.mul; return CountedF(a.v * b.v); }
inline CountedF operator/(CountedF a, CountedF b) { ++g_flops.div; return CountedF(a.v / b.v); }
inline bool operator<(CountedF a, CountedF b) { ++g_flops.cmp; return a.v < b.v; }
inline bool operator>(CountedF a, CountedF b) { ++g_flops.cmp; return a.v > b.v; }
inline bool operator<=(CountedF a, CountedF b) { ++g_flops.cmp; return a.v <= b.v; }
inline bool operator>=(CountedF a, CountedF b) { ++g_flops.cmp; return a.v >= b.v; }
inline bool operator==(CountedF a, CountedF b) { ++g_flops.cmp; return a.v == b.v; }
inline bool operator!=(CountedF a, CountedF b) { ++g_flops.cmp; return a.v != b.v; }

}  // namespace oracle

namespace std {
inline oracle::CountedF sqrt(oracle::CountedF a) { ++oracle::g_flops.sqrt; return oracle::CountedF(std::sqrt(a.v)); }
inline oracle::CountedF fabs(oracle::CountedF a) { ++oracle::g_flops.other; return oracle::CountedF(std::fabs(a.v)); }
inline oracle::CountedF fmin(oracle::CountedF a, oracle::CountedF b) {
    ++oracle::g_flops.other;
    return oracle::CountedF(std::fmin(a.v, b.v));
}
inline oracle::CountedF fmax(oracle::CountedF a, oracle::CountedF b) {
    ++oracle::g_flops.other;
    return oracle::CountedF(std::fmax(a.v, b.v));
}
inline oracle::CountedF copysign(oracle::CountedF a, oracle::CountedF b) {
    ++oracle::g_flops.other;
    return oracle::CountedF(std::copysign(a.v, b.v));
}
inline bool isinf(oracle::CountedF a) { return std::isinf(a.v); }
inline bool isfinite(oracle::CountedF a) { return std::isfinite(a.v); }
inline bool isnan(oracle::CountedF a) { return std::isnan(a.v); }
template <> struct numeric_limits<oracle::CountedF> {
    static constexpr bool is_specialized = true;
    static constexpr oracle::CountedF max() { return oracle::CountedF(numeric_limits<float>::max()); }
    static constexpr oracle::CountedF lowest() { return oracle::CountedF(numeric_limits<float>::lowest()); }
    static constexpr oracle::CountedF epsilon() { return oracle::CountedF(numeric_limits<float>::epsilon()); }
    static constexpr oracle::CountedF infinity() { return oracle::CountedF(numeric_limits<float>::infinity()); }
};
}  // namespace std

#include "parry_oracle.hpp"
#include "parry_oracle_contact.hpp"

namespace oracle {

template <> struct Consts<CountedF> {
    static constexpr CountedF bin_eps = CountedF(1.0e-5f);
    static constexpr CountedF max_value = CountedF(3.40282347e+38f);
};
template <> struct Eps<CountedF> {
    static constexpr CountedF eps = CountedF(1.1920929e-07f);
};

}  // namespace oracle

namespace {

using oracle::CountedF;

struct ShapeDescF {  // the 32-byte pq_shape_desc layout
    uint32_t type, first_vertex, num_vertices, reserved;
    float params[4];
};

}  // namespace

extern "C" {

// The oracle's query::contact (as orc_contact_pairs_f32, one thread) on every
// pair, counting its arithmetic.  counts[6]: add/sub, mul, div, sqrt (the flops),
// comparisons, min/max/abs/copysign; per_pair (nullable, m entries): the flops
// (add + mul + div + sqrt) of each pair; dist/status (nullable): the contact's
// dist and status, as orc_contact_pairs_f32 reports them.
void orc_contact_flops_f32(const void *descs, uint32_t nshapes, const float *verts, uint32_t nverts,
                           const float *poses, const uint32_t *pairs, uint64_t m, float prediction,
                           const float *init_dirs, uint64_t *counts, uint32_t *per_pair, float *dist,
                           int32_t *status) {
    using namespace oracle;
    const ShapeDescF *d = (const ShapeDescF *)descs;
    std::vector<V3<CountedF>> pts(nverts);
    for (uint32_t k = 0; k < nverts; ++k) pts[k] = v3(CountedF(verts[3 * k]), CountedF(verts[3 * k + 1]),
                                                      CountedF(verts[3 * k + 2]));
    std::vector<Shape<CountedF>> shapes(nshapes);
    for (uint32_t i = 0; i < nshapes; ++i) {
        shapes[i].type = d[i].type;
        shapes[i].pts = pts.data() + d[i].first_vertex;
        shapes[i].n = d[i].num_vertices;
        for (int k = 0; k < 4; ++k) shapes[i].p[k] = CountedF(d[i].params[k]);
    }
    g_flops = FlopCount{};
    for (uint64_t k = 0; k < m; ++k) {
        const uint64_t before = g_flops.add + g_flops.mul + g_flops.div + g_flops.sqrt;
        uint32_t i = pairs[2 * k], j = pairs[2 * k + 1];
        const float *a = poses + 7 * (size_t)i, *c = poses + 7 * (size_t)j;
        Iso<CountedF> p1{Quat<CountedF>{a[0], a[1], a[2], a[3]}, v3(CountedF(a[4]), CountedF(a[5]), CountedF(a[6]))};
        Iso<CountedF> p2{Quat<CountedF>{c[0], c[1], c[2], c[3]}, v3(CountedF(c[4]), CountedF(c[5]), CountedF(c[6]))};
        V3<CountedF> seed;
        const V3<CountedF> *sp = nullptr;
        if (init_dirs) {
            seed = v3(CountedF(init_dirs[3 * k]), CountedF(init_dirs[3 * k + 1]), CountedF(init_dirs[3 * k + 2]));
            sp = &seed;
        }
        Contact<CountedF> ct{};
        ContactStats st;
        const int r = contact(p1, shapes[i], p2, shapes[j], CountedF(prediction), ct, sp, &st);
        if (status) status[k] = r;
        if (dist) dist[k] = r == 1 ? ct.dist.v : 0.0f;
        if (per_pair) per_pair[k] = (uint32_t)(g_flops.add + g_flops.mul + g_flops.div + g_flops.sqrt - before);
    }
    counts[0] = g_flops.add;
    counts[1] = g_flops.mul;
    counts[2] = g_flops.div;
    counts[3] = g_flops.sqrt;
    counts[4] = g_flops.cmp;
    counts[5] = g_flops.other;
}

}  // extern "C"
