/*
 * Oracle (test infrastructure): 3-vector helpers with the reference's
 * operation order (src/vec3d.hpp, MUSCL_geometry.hpp:1154-1170).
 */
#ifndef SPLASH_ORACLE_VEC_H
#define SPLASH_ORACLE_VEC_H

#include <math.h>

typedef struct v3 {
    double x[3];
} v3;

static inline v3 v3m(double a, double b, double c)
{
    v3 r = {{a, b, c}};
    return r;
}
static inline v3 v3ld(const double *p) { return v3m(p[0], p[1], p[2]); }
static inline v3 v3add(v3 a, v3 b) { return v3m(a.x[0] + b.x[0], a.x[1] + b.x[1], a.x[2] + b.x[2]); }
static inline v3 v3sub(v3 a, v3 b) { return v3m(a.x[0] - b.x[0], a.x[1] - b.x[1], a.x[2] - b.x[2]); }
static inline v3 v3muls(v3 a, double s) { return v3m(a.x[0] * s, a.x[1] * s, a.x[2] * s); }
static inline v3 v3divs(v3 a, double s) { return v3m(a.x[0] / s, a.x[1] / s, a.x[2] / s); }
/* vec3d.hpp:77-85: sqrt(std::reduce(x*x)) = sqrt(((0+x0^2)+x1^2)+x2^2) */
static inline double v3norm(v3 a)
{
    double s = 0.0;
    s = s + a.x[0] * a.x[0];
    s = s + a.x[1] * a.x[1];
    s = s + a.x[2] * a.x[2];
    return sqrt(s);
}
/* MUSCL_geometry.hpp:1154-1162 */
static inline v3 v3cross(v3 a, v3 b)
{
    return v3m(a.x[1] * b.x[2] - a.x[2] * b.x[1], a.x[2] * b.x[0] - a.x[0] * b.x[2],
               a.x[0] * b.x[1] - a.x[1] * b.x[0]);
}
/* MUSCL_geometry.hpp:1164-1170 */
static inline double v3dot(v3 a, v3 b) { return a.x[0] * b.x[0] + a.x[1] * b.x[1] + a.x[2] * b.x[2]; }

/* MUSCL_geometry.hpp:981-987 great-circle distance */
static inline double gc_distance(v3 a, v3 b)
{
    a = v3divs(a, v3norm(a));
    b = v3divs(b, v3norm(b));
    double d = v3norm(v3sub(a, b));
    return 2 * asin(d / 2);
}

/* std::min / std::max of the C++ standard library: min(a,b) = b<a ? b : a,
 * max(a,b) = a<b ? b : a (NaN operands follow these comparisons). */
static inline double smin(double a, double b) { return (b < a) ? b : a; }
static inline double smax(double a, double b) { return (a < b) ? b : a; }

#endif
