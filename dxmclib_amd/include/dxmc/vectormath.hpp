// Small 3-vector helpers used on the host (source geometry) and restated on
// the device. Function names follow the reference (vectormath.hpp:24-223) so
// user code keeps compiling; the floating point operation order of rotate(),
// cross(), argmin3() and peturb() is the reference's, because the exposure
// geometry and every scatter direction must round identically on the host,
// in the oracle and in the kernels.
#pragma once
#include "dxmc/floating.hpp"
#include <cmath>
#include <cstdint>
#include <type_traits>

namespace dxmc {
namespace vectormath {

    template <Floating T>
    inline T lenght_sqr(const T v[3]) noexcept { return v[0] * v[0] + v[1] * v[1] + v[2] * v[2]; }

    template <Floating T>
    inline T lenght(const T v[3]) noexcept { return std::sqrt(lenght_sqr(v)); }

    template <Floating T>
    inline void normalize(T v[3]) noexcept
    {
        const T inv = T { 1 } / std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        v[0] *= inv;
        v[1] *= inv;
        v[2] *= inv;
    }

    template <Floating T>
    inline T dot(const T a[3], const T b[3]) noexcept { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

    // res = a x b
    template <Floating T>
    inline void cross(const T a[3], const T b[3], T res[3]) noexcept
    {
        res[0] = a[1] * b[2] - a[2] * b[1];
        res[1] = a[2] * b[0] - a[0] * b[2];
        res[2] = a[0] * b[1] - a[1] * b[0];
    }
    // res = c[0:3] x c[3:6] (beam direction from a pair of direction cosines)
    template <Floating T>
    inline void cross(const T c[6], T res[3]) noexcept { cross(c, c + 3, res); }

    // Rodrigues rotation of v about a unit axis by angle (radians).
    // angle between two vectors by the cancellation-free formula of
    // W. Kahan ("Mindless" (2006) section 12) on the triangle a, b, a - b
    template <Floating T>
    inline T angleBetween(const T a[3], const T b[3]) noexcept
    {
        const T d[3] = { a[0] - b[0], a[1] - b[1], a[2] - b[2] };
        const T la = lenght(a), lb = lenght(b), lc = lenght(d);
        const T mu = lb >= lc ? lc - (la - lb) : lb - (la - lc);
        return T { 2 } * std::atan(std::sqrt(((la - lb) + lc) * mu / ((la + (lb + lc)) * ((la - lc) + lb))));
    }
    // signed angle from v1 to v2 about the plane normal (all three are
    // normalised in place, as in the reference)
    template <Floating T>
    inline T angleBetweenOnPlane(T v1[3], T v2[3], T planeNormal[3]) noexcept
    {
        normalize(v1);
        normalize(v2);
        normalize(planeNormal);
        T c[3];
        cross(v1, v2, c);
        return std::atan2(dot(c, planeNormal), dot(v1, v2));
    }

    template <Floating T>
    inline void rotate(T v[3], const T axis[3], const T angle) noexcept
    {
        const T s = std::sin(angle);
        const T c = std::cos(angle);
        const T m = (T { 1 } - c) * dot(v, axis);
        const T r0 = c * v[0] + m * axis[0] + s * (axis[1] * v[2] - axis[2] * v[1]);
        const T r1 = c * v[1] + m * axis[1] + s * (-axis[0] * v[2] + axis[2] * v[0]);
        const T r2 = c * v[2] + m * axis[2] + s * (axis[0] * v[1] - axis[1] * v[0]);
        v[0] = r0;
        v[1] = r1;
        v[2] = r2;
    }

    template <Floating T>
    inline void projectToPlane(T v[3], const T n[3]) noexcept
    {
        const T d = dot(v, n);
        for (int i = 0; i < 3; ++i)
            v[i] -= d * n[i];
    }

    // index of the component with the smallest magnitude (ties -> lower index,
    // except x==z<y which resolves to x, as in the reference's ternary chain)
    template <typename U, Floating T>
    inline U argmin3(const T v[3]) noexcept
    {
        const T ax = std::abs(v[0]), ay = std::abs(v[1]), az = std::abs(v[2]);
        if (ax <= ay)
            return ax <= az ? U { 0 } : U { 2 };
        return ay <= az ? U { 1 } : U { 2 };
    }
    template <typename U, Floating T>
    inline U argmax3(const T v[3]) noexcept
    {
        const T ax = std::abs(v[0]), ay = std::abs(v[1]), az = std::abs(v[2]);
        if (ax >= ay)
            return ax >= az ? U { 0 } : U { 2 };
        return ay >= az ? U { 1 } : U { 2 };
    }

    // express v in the basis (b1,b2,b3) given in world coordinates
    template <Floating T>
    inline void changeBasisInverse(const T b1[3], const T b2[3], const T b3[3], T v[3]) noexcept
    {
        const T n0 = b1[0] * v[0] + b1[1] * v[1] + b1[2] * v[2];
        const T n1 = b2[0] * v[0] + b2[1] * v[1] + b2[2] * v[2];
        const T n2 = b3[0] * v[0] + b3[1] * v[1] + b3[2] * v[2];
        v[0] = n0;
        v[1] = n1;
        v[2] = n2;
    }
    template <Floating T>
    inline void changeBasis(const T b1[3], const T b2[3], const T b3[3], T v[3]) noexcept
    {
        const T n0 = b1[0] * v[0] + b2[0] * v[1] + b3[0] * v[2];
        const T n1 = b1[1] * v[0] + b2[1] * v[1] + b3[1] * v[2];
        const T n2 = b1[2] * v[0] + b2[2] * v[1] + b3[2] * v[2];
        v[0] = n0;
        v[1] = n1;
        v[2] = n2;
    }

    // Deflect a direction by polar angle theta and azimuth phi. The auxiliary
    // axis is dir x e_k (k = smallest |component|) and is deliberately NOT
    // normalised: the reference does the same (vectormath.hpp:197-221), so
    // |dir| shrinks after each deflection and so do the Woodcock steps.
    template <Floating T>
    inline void peturb(T dir[3], const T theta, const T phi) noexcept
    {
        T e[3] = { 0, 0, 0 };
        e[argmin3<std::uint32_t, T>(dir)] = T { 1 };
        T orth[3];
        cross(dir, e, orth);
        rotate(orth, dir, phi);
        const T st = std::sin(theta);
        const T ct = std::cos(theta);
        dir[0] = dir[0] * ct + orth[0] * st;
        dir[1] = dir[1] * ct + orth[1] * st;
        dir[2] = dir[2] * ct + orth[2] * st;
    }
}
}
