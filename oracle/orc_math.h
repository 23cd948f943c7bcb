// Oracle: small fixed-size vector/matrix helpers for the CPU restatement.
//
// TEST INFRASTRUCTURE ONLY. Nothing under oracle/ is part of the product; it is the
// checker that tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use.
//
// Every helper fixes its floating-point evaluation order explicitly (left to right,
// no contraction: the library is built with -ffp-contract=off) so that FP32 results
// are reproducible and comparable element by element with the GPU path.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>

namespace orc
{

template <class T>
struct V3
{
    T x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(T a, T b, T c) : x(a), y(b), z(c) {}
    T&       operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    const T& operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T> inline V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T> inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline V3<T> operator/(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline V3<T> cross(V3<T> a, V3<T> b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T> inline T length(V3<T> a) { return std::sqrt(dot(a, a)); }
// normalize = v * (1 / |v|)
template <class T> inline V3<T> normalize(V3<T> a) { return a * (T(1) / length(a)); }
template <class T> inline V3<T> vmin(V3<T> a, V3<T> b) { return {std::min(a.x, b.x), std::min(a.y, b.y), std::min(a.z, b.z)}; }
template <class T> inline V3<T> vmax(V3<T> a, V3<T> b) { return {std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z)}; }

// Row-major 3x3: m[r][c].
template <class T>
struct M3
{
    T m[3][3];
    M3() { for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) m[r][c] = 0; }
    static M3 diag(T d) { M3 a; a.m[0][0] = a.m[1][1] = a.m[2][2] = d; return a; }
};
template <class T> inline M3<T> operator+(const M3<T>& a, const M3<T>& b)
{
    M3<T> o; for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) o.m[r][c] = a.m[r][c] + b.m[r][c]; return o;
}
template <class T> inline M3<T> operator*(T s, const M3<T>& a)
{
    M3<T> o; for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) o.m[r][c] = s * a.m[r][c]; return o;
}
// y = M x, each row summed left to right
template <class T> inline V3<T> operator*(const M3<T>& a, V3<T> x)
{
    return {a.m[0][0] * x.x + a.m[0][1] * x.y + a.m[0][2] * x.z,
            a.m[1][0] * x.x + a.m[1][1] * x.y + a.m[1][2] * x.z,
            a.m[2][0] * x.x + a.m[2][1] * x.y + a.m[2][2] * x.z};
}
// outer(a, b)[r][c] = a_r * b_c
template <class T> inline M3<T> outer(V3<T> a, V3<T> b)
{
    M3<T> o; for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) o.m[r][c] = a[r] * b[c]; return o;
}
// General 3x3 inverse by the adjugate divided by the determinant (cofactor expansion
// along the first row). Used for the block-Jacobi preconditioner.
template <class T> inline M3<T> inverse(const M3<T>& a)
{
    const T(&m)[3][3] = a.m;
    T c00 = m[1][1] * m[2][2] - m[1][2] * m[2][1];
    T c01 = m[1][2] * m[2][0] - m[1][0] * m[2][2];
    T c02 = m[1][0] * m[2][1] - m[1][1] * m[2][0];
    T det = m[0][0] * c00 + m[0][1] * c01 + m[0][2] * c02;
    T inv = T(1) / det;
    M3<T> o;
    o.m[0][0] = c00 * inv;
    o.m[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) * inv;
    o.m[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * inv;
    o.m[1][0] = c01 * inv;
    o.m[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * inv;
    o.m[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) * inv;
    o.m[2][0] = c02 * inv;
    o.m[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) * inv;
    o.m[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) * inv;
    return o;
}

template <class T> inline T clampv(T v, T lo, T hi) { return std::max(std::min(v, hi), lo); }

}  // namespace orc
