/*
 * cfo_vec.h -- oracle 3-vector helpers with the reference Vector3 semantics
 * (src/common/include/cf/common/Vector.h): every operation is component-wise
 * and left-to-right, dot = (ax*bx + ay*by) + az*bz, and `s * v` evaluates as
 * `v * s`.  TEST INFRASTRUCTURE ONLY.
 */
#ifndef CFO_VEC_H
#define CFO_VEC_H

#include <math.h>

typedef struct { double x, y, z; } V3;

static inline V3 v3(double x, double y, double z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }
static inline V3 vzero(void) { return v3(0.0, 0.0, 0.0); }
static inline V3 vload(const double* p) { return v3(p[0], p[1], p[2]); }
static inline void vstore(double* p, V3 a) { p[0] = a.x; p[1] = a.y; p[2] = a.z; }
static inline V3 vadd(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V3 vsub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V3 vneg(V3 a) { return v3(-a.x, -a.y, -a.z); }
static inline V3 vmul(V3 a, double s) { return v3(a.x * s, a.y * s, a.z * s); }
static inline V3 vdiv(V3 a, double s) { return v3(a.x / s, a.y / s, a.z / s); }
static inline double vdot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline V3 vcross(V3 a, V3 b)
{
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline double vmag(V3 a) { return sqrt(vdot(a, a)); }
static inline int veq(V3 a, V3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
static inline double vget(V3 a, int c) { return c == 0 ? a.x : (c == 1 ? a.y : a.z); }

#endif
