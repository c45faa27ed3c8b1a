// commondefs.h -- value types of the drop-in C++ layer.  Same names and
// meaning as the reference's (include/commondefs.h:13-85, which takes them
// from trimesh2: point, ivec2/3/4, uTriFace, xform, eulerchar).
#pragma once

#include <array>
#include <cstdint>
#include <cstdlib>
#include <iostream>

template <int N, class T>
struct VcVec
{
    std::array<T, N> v{};
    VcVec() = default;
    explicit VcVec(T a)
    {
        v.fill(a);
    }
    template <class... A, class = typename std::enable_if<sizeof...(A) == N>::type>
    VcVec(A... a) : v{{static_cast<T>(a)...}}
    {
    }
    T& operator[](int i) { return v[i]; }
    const T& operator[](int i) const { return v[i]; }
    bool operator==(const VcVec& o) const { return v == o.v; }
    bool operator!=(const VcVec& o) const { return v != o.v; }
    bool operator<(const VcVec& o) const { return v < o.v; }
    const T* data() const { return v.data(); }
    T* begin() { return v.data(); }
    T* end() { return v.data() + N; }
    const T* begin() const { return v.data(); }
    const T* end() const { return v.data() + N; }
    static constexpr int size() { return N; }
};

using point = VcVec<3, float>;
using ivec2 = VcVec<2, int>;
using ivec3 = VcVec<3, int>;
using ivec4 = VcVec<4, int>;
using uTriFace = VcVec<3, unsigned int>;

// 4x4 voxel <-> model transform, column-major doubles as trimesh::xform
// (dense volumes and .sof octrees use the identity; .sog octrees a scale and
// a translation, src/octree.cpp:461-475)
struct xform
{
    std::array<double, 16> m{{1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1}};
    static xform identity() { return xform(); }
    static xform trans(double x, double y, double z)
    {
        xform t;
        t.m[12] = x;
        t.m[13] = y;
        t.m[14] = z;
        return t;
    }
    static xform trans(const VcVec<3, float>& p) { return trans(p[0], p[1], p[2]); }
    static xform scale(double s)
    {
        xform t;
        t.m[0] = t.m[5] = t.m[10] = s;
        return t;
    }
    double& operator[](int i) { return m[i]; }
    double operator[](int i) const { return m[i]; }
    const double* data() const { return m.data(); }
};

// matrix product, each entry summed left to right over k
inline xform operator*(const xform& a, const xform& b)
{
    xform c;
    for (int j = 0; j < 4; ++j)
        for (int i = 0; i < 4; ++i)
        {
            double s = a[i] * b[4 * j];
            for (int k = 1; k < 4; ++k)
                s += a[i + 4 * k] * b[k + 4 * j];
            c.m[i + 4 * j] = s;
        }
    return c;
}

extern "C" int vc_xform_apply(const double* m, const float* pts, int64_t n,
                              float* out);

// point transform (homogeneous, double accumulation, rounded to float)
inline VcVec<3, float> operator*(const xform& x, const VcVec<3, float>& p)
{
    VcVec<3, float> q;
    vc_xform_apply(x.data(), p.data(), 1, q.begin());
    return q;
}

// Euler characteristic summary (reference include/commondefs.h:60-85)
struct eulerchar
{
    int V = 0, E = 0, F = 0, C = 0, T = 0, euler = 0;
    eulerchar() = default;
    eulerchar(int _v, int _e, int _f, int _t, int _c, int _euler)
        : V(_v), E(_e), F(_f), C(_c), T(_t), euler(_euler)
    {
    }
    void logToConsole(const char* name) const
    {
        std::cout << "------ Euler Characteristic (of " << name << ") -------\n"
                  << "euler = V - E + F = " << euler << "\n"
                  << "V / E / F / T / C = " << V << " / " << E << " / " << F
                  << " / " << T << " / " << C << std::endl;
    }
};
