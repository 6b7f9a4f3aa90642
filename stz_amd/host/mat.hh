// mat.hh -- 2x2 matrix value type with the interface the multigrid facade
// needs from the reference's shared/mat.hh (a,b,c,d, products, inverse).
#ifndef STZ_AMD_MAT_HH
#define STZ_AMD_MAT_HH

#include "vec.hh"

struct mat {
    double a, b, c, d;
    mat() {}
    mat(double a_) : a(a_), b(0), c(0), d(a_) {}
    mat(double a_, double b_, double c_, double d_) : a(a_), b(b_), c(c_), d(d_) {}
    mat operator+(mat p) const { return mat(a + p.a, b + p.b, c + p.c, d + p.d); }
    mat operator-(mat p) const { return mat(a - p.a, b - p.b, c - p.c, d - p.d); }
    mat operator*(double e) const { return mat(a * e, b * e, c * e, d * e); }
    mat operator*(mat e) const {
        return mat(a * e.a + b * e.c, a * e.b + b * e.d, c * e.a + d * e.c, c * e.b + d * e.d);
    }
    vec operator*(vec e) const { return vec(a * e.x + b * e.y, c * e.x + d * e.y); }
    double det() const { return a * d - b * c; }
    mat inverse() const {
        double idet = 1.0 / det();
        return mat(idet * d, -idet * b, -idet * c, idet * a);
    }
    mat transpose() const { return mat(a, c, b, d); }
    double mod_sq() const { return a * a + b * b + c * c + d * d; }
};
inline double mod_sq(mat m) { return m.mod_sq(); }

#endif
