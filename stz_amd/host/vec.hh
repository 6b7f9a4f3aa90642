// vec.hh -- 2-vector value type with the interface of the reference's
// shared/vec.hh (x, y, arithmetic, |v|^2 conversion) for host code using the
// stz_amd facade.
#ifndef STZ_AMD_VEC_HH
#define STZ_AMD_VEC_HH

#include <cmath>

struct vec {
    double x, y;
    vec() {}
    vec(double x_) : x(x_), y(x_) {}
    vec(double x_, double y_) : x(x_), y(y_) {}
    vec operator+(vec p) const { return vec(x + p.x, y + p.y); }
    vec operator-(vec p) const { return vec(x - p.x, y - p.y); }
    vec operator*(double a) const { return vec(x * a, y * a); }
    vec operator/(double a) const { return vec(x / a, y / a); }
    void operator+=(vec p) { x += p.x; y += p.y; }
    void operator-=(vec p) { x -= p.x; y -= p.y; }
    void operator*=(double a) { x *= a; y *= a; }
    operator double() const { return x * x + y * y; }
    double magnitude() const { return std::sqrt(x * x + y * y); }
};
inline vec operator*(double e, vec f) { return vec(e * f.x, e * f.y); }
inline vec operator-(vec f) { return vec(-f.x, -f.y); }
inline double mod_sq(vec a) { return a.x * a.x + a.y * a.y; }

#endif
