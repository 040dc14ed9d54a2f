// parry_oracle_cast.hpp -- TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of parry3d's linear shape casting (time of impact under
// translational motion) for the GPU parity tests: query::cast_shapes
// (src/query/shape_cast/shape_cast.rs:274-286) through
// DefaultQueryDispatcher::cast_shapes (default_query_dispatcher.rs:434-540) for
// the shapes held here -- Ball-Ball in closed form
// (shape_cast_ball_ball.rs:10-68, ray_ball.rs:33-77) and every other pair as
// support maps (shape_cast_support_map_support_map.rs:11-72): GJK's ray cast on
// the Minkowski difference (gjk::directional_distance and minkowski_ray_cast,
// gjk.rs:632-795), the rounded first shape when target_distance > 0
// (RoundShapeRef, round_shape.rs:317-343), and contact_support_map_support_map
// for the impact geometry on penetration.  Same conventions as
// parry_oracle.hpp: T = float is parry3d, T = double parry3d-f64, and every
// expression keeps the reference's operation order.
#pragma once

#include <cmath>
#include <cstdint>

#include "parry_oracle.hpp"
#include "parry_oracle_contact.hpp"

namespace oracle {

// ShapeCastOptions (shape_cast.rs:105-146; default: max_toi Real::MAX, target 0,
// stop_at_penetration and compute_impact_geometry_on_penetration true).
template <class T> struct CastOptions {
    T max_toi;
    T target_distance;
    bool stop_at_penetration;
    bool compute_impact_geometry_on_penetration;
};

// ShapeCastStatus (shape_cast.rs:12-33), numbered from 1; 0 is None.
enum CastStatus : uint32_t {
    CAST_NONE = 0,
    CAST_OUT_OF_ITERATIONS = 1,
    CAST_CONVERGED = 2,
    CAST_FAILED = 3,
    CAST_PENETRATING = 4,
};

// ShapeCastHit (shape_cast.rs:36-70), local frames of the two shapes.
template <class T> struct CastHit {
    T toi;
    V3<T> witness1, witness2, normal1, normal2;
    uint32_t status;
};

// SupportMap::local_support_point_toward (support_map.rs:302-304) for a unit
// direction: Ball and Capsule take it as given (ball.rs:247-249,
// capsule.rs:412-418); the others are their local_support_point.
template <class T> V3<T> local_support_toward(const Shape<T> &s, V3<T> u) {
    if (s.type == BALL) return u * s.p[0];
    if (s.type == CAPSULE) {
        V3<T> a = v3(T(0), -s.p[0], T(0)), b = v3(T(0), s.p[0], T(0));
        if (dot(u, a) > dot(u, b)) return a + u * s.p[1];
        return b + u * s.p[1];
    }
    return local_support_point(s, u);
}

// CSOPoint::from_shapes (cso_point.rs:69-78) with shape 1 rounded by `border`
// when border > 0: RoundShapeRef::local_support_point normalizes the
// direction and adds it times the border (round_shape.rs:330-343).
template <class T>
CSO<T> cso_rounded(const Iso<T> &pos12, const Shape<T> &g1, T border, const Shape<T> &g2, V3<T> dir) {
    V3<T> sp1;
    if (border > T(0)) {
        V3<T> u = normalize(dir);
        sp1 = local_support_toward(g1, u) + u * border;
    } else {
        sp1 = local_support_point(g1, dir);
    }
    V3<T> sp2 = support_point(pos12, g2, -dir);
    return CSO<T>{sp1 - sp2, sp1, sp2};
}

template <class T> inline CSO<T> cso_translate(const CSO<T> &c, V3<T> d) { return CSO<T>{c.point + d, c.orig1, c.orig2}; }

// minkowski_ray_cast (gjk.rs:660-795) of a ray from the origin along `ray_dir`
// with max_time_of_impact = Real::MAX, without the fixed-point feature.
// Returns false for None; iters (nullable) counts the loop's passes.
template <class T>
bool minkowski_ray_cast(const Iso<T> &pos12, const Shape<T> &g1, T border, const Shape<T> &g2, V3<T> ray_dir,
                        Simplex<T> &simplex, T &toi, V3<T> &normal, uint32_t *iters) {
    const T eps_tol = Eps<T>::eps * T(10);
    const T eps_rel = std::sqrt(eps_tol);
    const T max_toi = Consts<T>::max_value;
    const T ray_length = norm(ray_dir);
    if (iters) *iters = 0;
    if (rel_eq(ray_length, T(0))) return false;
    T ltoi = T(0);
    V3<T> origin = v3(T(0), T(0), T(0));
    const V3<T> cdir = ray_dir / ray_length;
    V3<T> dir = -cdir;
    V3<T> ldir = dir;
    simplex.reset(cso_translate(cso_rounded(pos12, g1, border, g2, dir), -origin));
    V3<T> proj = simplex.project_origin_and_reduce();
    T max_bound = Consts<T>::max_value;
    uint32_t niter = 0;
    bool last_chance = false;
    for (;;) {
        const T old_max_bound = max_bound;
        V3<T> u;
        T dist;
        if (try_unit(-proj, eps_tol, u, dist)) {
            dir = u;
            max_bound = dist;
        } else {
            toi = ltoi / ray_length;
            normal = ldir;
            return true;
        }
        CSO<T> sp;
        if (max_bound >= old_max_bound) {
            last_chance = true;  // CSOPoint::single_point (cso_point.rs:42-44)
            V3<T> p = proj + origin;
            sp = CSO<T>{p, p, v3(T(0), T(0), T(0))};
        } else {
            sp = cso_rounded(pos12, g1, border, g2, dir);
        }
        if (last_chance && ltoi > T(0)) {
            toi = ltoi / ray_length;
            normal = ldir;
            return true;
        }
        // ray_toi_with_halfspace(support point, dir, current ray) (ray_halfspace.rs:9-39)
        bool have_t = false;
        T t = T(0);
        {
            V3<T> dpos = sp.point - origin;
            T denom = dot(dir, cdir);
            if (!rel_eq(denom, T(0))) {
                t = dot(dir, dpos) / denom;
                have_t = t >= T(0);
            }
        }
        if (have_t) {
            if (dot(dir, cdir) < T(0) && t > T(0)) {
                ldir = dir;
                ltoi += t;
                if (ltoi / ray_length > max_toi) return false;
                V3<T> shift = cdir * t;
                origin = origin + shift;
                max_bound = Consts<T>::max_value;
                for (size_t i = 0; i < simplex.dim + 1; ++i) simplex.vertices[i].point = simplex.vertices[i].point + (-shift);
                last_chance = false;
            }
        } else if (dot(dir, cdir) > eps_tol) {
            return false;
        }
        if (last_chance) return false;
        const T min_bound = -dot(dir, sp.point - origin);
        if (max_bound - min_bound <= eps_rel * max_bound) return false;
        simplex.add_point(cso_translate(sp, -origin));
        proj = simplex.project_origin_and_reduce();
        if (simplex.dim == 3) {
            if (min_bound >= eps_tol) return false;
            toi = ltoi / ray_length;
            normal = ldir;
            return true;
        }
        niter += 1;
        if (iters) *iters = niter;
        if (niter == 100) return false;
    }
}

// cast_shapes_support_map_support_map (shape_cast_support_map_support_map.rs:11-72),
// with gjk::directional_distance (gjk.rs:632-657) inlined.
template <class T>
bool cast_sm_sm(const Iso<T> &pos12, V3<T> vel12, const Shape<T> &g1, const Shape<T> &g2, const CastOptions<T> &o,
                CastHit<T> &hit, uint32_t *iters) {
    Simplex<T> simplex;
    T toi;
    V3<T> n1;
    const T border = o.target_distance > T(0) ? o.target_distance : T(0);
    if (!minkowski_ray_cast(pos12, g1, border, g2, vel12, simplex, toi, n1, iters)) return false;
    V3<T> w1 = v3(T(0), T(0), T(0)), w2 = v3(T(0), T(0), T(0));
    if (toi != T(0)) gjk_result(simplex, simplex.dim == 3, w1, w2);
    if (toi > o.max_toi) return false;
    if ((o.compute_impact_geometry_on_penetration || !o.stop_at_penetration) && toi < T(1.0e-5)) {
        Contact<T> c;
        if (!contact_sm_sm(pos12, g1, g2, Consts<T>::max_value, c, static_cast<const V3<T> *>(nullptr), nullptr))
            return false;
        const T normal_vel = dot(c.normal1, vel12);
        if (!o.stop_at_penetration && normal_vel >= T(0)) return false;
        hit = CastHit<T>{toi, c.point1, c.point2, c.normal1, c.normal2, CAST_PENETRATING};
        return true;
    }
    hit.toi = toi;
    hit.normal1 = n1;
    hit.normal2 = inv_rotate(pos12.q, -n1);
    hit.witness1 = w1 - n1 * o.target_distance;
    hit.witness2 = inv_rotate(pos12.q, w2 - pos12.t);
    hit.status = toi == T(0) ? CAST_PENETRATING : CAST_CONVERGED;
    return true;
}

// cast_shapes_ball_ball (shape_cast_ball_ball.rs:10-68) with ray_toi_with_ball,
// solid = true (ray_ball.rs:33-77).
template <class T>
bool cast_ball_ball(const Iso<T> &pos12, V3<T> vel12, T r1, T r2, const CastOptions<T> &o, CastHit<T> &hit) {
    const T rsum = r1 + r2 + o.target_distance;
    const T radius = rsum;
    const V3<T> center = -pos12.t;
    const V3<T> origin = v3(T(0), T(0), T(0));
    bool inside = false;
    T toi;
    {
        const V3<T> dcenter = origin - center;
        const T a = dot(vel12, vel12);
        const T b = dot(dcenter, vel12);
        const T c = dot(dcenter, dcenter) - radius * radius;
        if (a == T(0)) {
            if (c > T(0)) return false;
            inside = true;
            toi = T(0);
        } else if (c > T(0) && b > T(0)) {
            return false;
        } else {
            const T delta = b * b - a * c;
            if (delta < T(0)) return false;
            const T t = (-b - std::sqrt(delta)) / a;
            if (t <= T(0)) {
                inside = true;
                toi = T(0);
            } else {
                toi = t;
            }
        }
    }
    if (toi > o.max_toi) return false;
    const V3<T> dpt = (origin + vel12 * toi) - center;
    V3<T> n1, n2, w1, w2;
    if (radius == T(0)) {
        n1 = v3(T(1), T(0), T(0));
        n2 = inv_rotate(pos12.q, -v3(T(1), T(0), T(0)));
        w1 = w2 = v3(T(0), T(0), T(0));
    } else {
        n1 = dpt / radius;
        n2 = inv_rotate(pos12.q, -n1);
        w1 = n1 * r1;
        w2 = n2 * r2;
    }
    if (!o.stop_at_penetration && toi < T(1.0e-5) && dot(n1, vel12) >= T(0)) return false;
    hit.toi = toi;
    hit.normal1 = n1;
    hit.normal2 = n2;
    hit.witness1 = w1;
    hit.witness2 = w2;
    hit.status = inside && dot(center, center) < rsum * rsum ? CAST_PENETRATING : CAST_CONVERGED;
    return true;
}

// query::cast_shapes (shape_cast.rs:274-286): pos12 = pos1^-1 pos2 and
// vel12 = pos1^-1 (vel2 - vel1), then the dispatch.  Returns false for None.
template <class T>
bool cast_shapes(const Iso<T> &pos1, V3<T> vel1, const Shape<T> &g1, const Iso<T> &pos2, V3<T> vel2,
                 const Shape<T> &g2, const CastOptions<T> &o, CastHit<T> &hit, uint32_t *iters) {
    const Iso<T> pos12 = inv_mul(pos1, pos2);
    const V3<T> vel12 = inv_rotate(pos1.q, vel2 - vel1);
    if (iters) *iters = 0;
    if (g1.type == BALL && g2.type == BALL) return cast_ball_ball(pos12, vel12, g1.p[0], g2.p[0], o, hit);
    return cast_sm_sm(pos12, vel12, g1, g2, o, hit, iters);
}

}  // namespace oracle
