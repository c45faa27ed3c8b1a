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

// affine voxel <-> model transform; the reference's Surfacer and thinning use
// the identity (src/surfacing.cpp and include/surfacing.h:97-119)
struct xform
{
    static xform identity() { return xform(); }
};

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
