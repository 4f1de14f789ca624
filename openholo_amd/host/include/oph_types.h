// Basic types of the ophPointCloud drop-in: the same names and layouts the
// reference's public headers expose to callers (Openholo/src/typedef.h:49-57,
// struct.h:49-106, define.h:155-159, vec.h / ivec.h, complex.h:84-123),
// written here as small self-contained definitions.
#pragma once

#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

namespace oph {

typedef double Real;
typedef unsigned int uint;
typedef unsigned char uchar;
typedef unsigned long long ulonglong;
typedef long long longlong;

#ifndef _X
#define _X 0
#endif
#ifndef _Y
#define _Y 1
#endif
#ifndef _Z
#define _Z 2
#endif
#ifndef _RE
#define _RE 0
#endif
#ifndef _IM
#define _IM 1
#endif

// runtime mode bits (Openholo/src/define.h:155-159)
#define MODE_CPU 0
#define MODE_GPU 1
#define MODE_DOUBLE 0
#define MODE_FLOAT 2
#define MODE_FASTMATH 4
// matrix-core Fresnel variants (extensions of this library; RS ignores them)
#define MODE_MFMA_BF16X1 0x10
#define MODE_MFMA_BF16X3 0x20
#define MODE_MFMA_FP16X3 0x40

#define OPH_LOG(...) std::printf(__VA_ARGS__)

struct vec2 {
    Real v[2];
    vec2() : v{0, 0} {}
    vec2(Real a, Real b) : v{a, b} {}
    Real& operator[](int i) { return v[i]; }
    const Real& operator[](int i) const { return v[i]; }
};
struct vec3 {
    Real v[3];
    vec3() : v{0, 0, 0} {}
    vec3(Real a, Real b, Real c) : v{a, b, c} {}
    Real& operator[](int i) { return v[i]; }
    const Real& operator[](int i) const { return v[i]; }
};
struct ivec2 {
    int v[2];
    ivec2() : v{0, 0} {}
    ivec2(int a, int b) : v{a, b} {}
    int& operator[](int i) { return v[i]; }
    const int& operator[](int i) const { return v[i]; }
};

// std::complex with the reference's element access and helpers
template <typename T = double>
class Complex : public std::complex<T> {
public:
    Complex() : std::complex<T>(0, 0) {}
    Complex(T re, T im) : std::complex<T>(re, im) {}
    Complex(const std::complex<T>& c) : std::complex<T>(c) {}
    T& operator[](int i) { return reinterpret_cast<T(&)[2]>(*this)[i]; }
    const T& operator[](int i) const { return reinterpret_cast<const T(&)[2]>(*this)[i]; }
    T mag2() const { return this->real() * this->real() + this->imag() * this->imag(); }
    T mag() const { return std::sqrt(mag2()); }
    T angle() const { return std::atan2(this->imag(), this->real()); }
};

#pragma pack(push, 1)
struct Point { Real pos[3]; };
struct Color { Real color[3]; };
struct Vertex {
    Point point;
    Color color;
    Real phase;
};
#pragma pack(pop)
static_assert(sizeof(Vertex) == 56, "Vertex must keep the reference's 56-byte layout");

enum FLIP { NONE = 0, VERTICAL = 1, HORIZONTAL = 2, BOTH = 3 };

}  // namespace oph

using namespace oph;
