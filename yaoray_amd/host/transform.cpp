// Column-major 4x4 transforms (translation in m[12..14]) with the reference's
// operation order, so instance matrices and their inverses are bit-identical
// to the ones the reference uploads (src/core/transform.cpp:14-169).
#include "yr_ir.hpp"

#include <cmath>

namespace yr {

Mat4f Multiply(Mat4f a, Mat4f b) {
    Mat4f r;
    r.m.fill(0.0f);
    for (int col = 0; col < 4; ++col) {
        for (int row = 0; row < 4; ++row) {
            for (int k = 0; k < 4; ++k) {
                r.m[static_cast<size_t>(col * 4 + row)] +=
                    a.m[static_cast<size_t>(k * 4 + row)] * b.m[static_cast<size_t>(col * 4 + k)];
            }
        }
    }
    return r;
}

Point3f TransformPoint(Mat4f t, Point3f p) {
    return Point3f{t.m[0] * p.x + t.m[4] * p.y + t.m[8] * p.z + t.m[12],
                   t.m[1] * p.x + t.m[5] * p.y + t.m[9] * p.z + t.m[13],
                   t.m[2] * p.x + t.m[6] * p.y + t.m[10] * p.z + t.m[14]};
}

Vec3f TransformVector(Mat4f t, Vec3f v) {
    return Vec3f{t.m[0] * v.x + t.m[4] * v.y + t.m[8] * v.z, t.m[1] * v.x + t.m[5] * v.y + t.m[9] * v.z,
                 t.m[2] * v.x + t.m[6] * v.y + t.m[10] * v.z};
}

Mat4f TranslationMatrix(Vec3f t) {
    Mat4f r;
    r.m[12] = t.x;
    r.m[13] = t.y;
    r.m[14] = t.z;
    return r;
}

Mat4f ScaleMatrix(Vec3f s) {
    Mat4f r;
    r.m[0] = s.x;
    r.m[5] = s.y;
    r.m[10] = s.z;
    return r;
}

Mat4f RotationAxisMatrix(float angle_degrees, Vec3f axis) {
    const Vec3f u = Normalize(axis);
    if (LengthSquared(u) == 0.0f) return Mat4f{};
    const float rad = angle_degrees * 3.14159265358979323846f / 180.0f;
    const float c = std::cos(rad);
    const float s = std::sin(rad);
    const float t = 1.0f - c;
    Mat4f r;
    r.m[0] = t * u.x * u.x + c;
    r.m[1] = t * u.x * u.y + s * u.z;
    r.m[2] = t * u.x * u.z - s * u.y;
    r.m[4] = t * u.x * u.y - s * u.z;
    r.m[5] = t * u.y * u.y + c;
    r.m[6] = t * u.y * u.z + s * u.x;
    r.m[8] = t * u.x * u.z + s * u.y;
    r.m[9] = t * u.y * u.z - s * u.x;
    r.m[10] = t * u.z * u.z + c;
    return r;
}

namespace {

// Cofactor expansion of the classic adjugate 4x4 inverse (the one in MESA's
// gluInvertMatrix, also used by the reference). Entry k of the adjugate is
// the left-to-right sum of six signed triple products a[i]*a[j]*a[l]; the
// table lists (sign, i, j, l) in evaluation order so the rounding sequence is
// the reference's.
struct Term {
    signed char sign;
    unsigned char i, j, l;
};

constexpr Term kAdjugate[16][6] = {
    /* 0 */ {{+1, 5, 10, 15}, {-1, 5, 11, 14}, {-1, 9, 6, 15}, {+1, 9, 7, 14}, {+1, 13, 6, 11}, {-1, 13, 7, 10}},
    /* 1 */ {{-1, 1, 10, 15}, {+1, 1, 11, 14}, {+1, 9, 2, 15}, {-1, 9, 3, 14}, {-1, 13, 2, 11}, {+1, 13, 3, 10}},
    /* 2 */ {{+1, 1, 6, 15}, {-1, 1, 7, 14}, {-1, 5, 2, 15}, {+1, 5, 3, 14}, {+1, 13, 2, 7}, {-1, 13, 3, 6}},
    /* 3 */ {{-1, 1, 6, 11}, {+1, 1, 7, 10}, {+1, 5, 2, 11}, {-1, 5, 3, 10}, {-1, 9, 2, 7}, {+1, 9, 3, 6}},
    /* 4 */ {{-1, 4, 10, 15}, {+1, 4, 11, 14}, {+1, 8, 6, 15}, {-1, 8, 7, 14}, {-1, 12, 6, 11}, {+1, 12, 7, 10}},
    /* 5 */ {{+1, 0, 10, 15}, {-1, 0, 11, 14}, {-1, 8, 2, 15}, {+1, 8, 3, 14}, {+1, 12, 2, 11}, {-1, 12, 3, 10}},
    /* 6 */ {{-1, 0, 6, 15}, {+1, 0, 7, 14}, {+1, 4, 2, 15}, {-1, 4, 3, 14}, {-1, 12, 2, 7}, {+1, 12, 3, 6}},
    /* 7 */ {{+1, 0, 6, 11}, {-1, 0, 7, 10}, {-1, 4, 2, 11}, {+1, 4, 3, 10}, {+1, 8, 2, 7}, {-1, 8, 3, 6}},
    /* 8 */ {{+1, 4, 9, 15}, {-1, 4, 11, 13}, {-1, 8, 5, 15}, {+1, 8, 7, 13}, {+1, 12, 5, 11}, {-1, 12, 7, 9}},
    /* 9 */ {{-1, 0, 9, 15}, {+1, 0, 11, 13}, {+1, 8, 1, 15}, {-1, 8, 3, 13}, {-1, 12, 1, 11}, {+1, 12, 3, 9}},
    /* 10 */ {{+1, 0, 5, 15}, {-1, 0, 7, 13}, {-1, 4, 1, 15}, {+1, 4, 3, 13}, {+1, 12, 1, 7}, {-1, 12, 3, 5}},
    /* 11 */ {{-1, 0, 5, 11}, {+1, 0, 7, 9}, {+1, 4, 1, 11}, {-1, 4, 3, 9}, {-1, 8, 1, 7}, {+1, 8, 3, 5}},
    /* 12 */ {{-1, 4, 9, 14}, {+1, 4, 10, 13}, {+1, 8, 5, 14}, {-1, 8, 6, 13}, {-1, 12, 5, 10}, {+1, 12, 6, 9}},
    /* 13 */ {{+1, 0, 9, 14}, {-1, 0, 10, 13}, {-1, 8, 1, 14}, {+1, 8, 2, 13}, {+1, 12, 1, 10}, {-1, 12, 2, 9}},
    /* 14 */ {{-1, 0, 5, 14}, {+1, 0, 6, 13}, {+1, 4, 1, 14}, {-1, 4, 2, 13}, {-1, 12, 1, 6}, {+1, 12, 2, 5}},
    /* 15 */ {{+1, 0, 5, 10}, {-1, 0, 6, 9}, {-1, 4, 1, 10}, {+1, 4, 2, 9}, {+1, 8, 1, 6}, {-1, 8, 2, 5}},
};

}  // namespace

Mat4f Inverse(Mat4f m) {
    const float* a = m.m.data();
    Mat4f r;
    for (int k = 0; k < 16; ++k) {
        float acc = 0.0f;
        for (int n = 0; n < 6; ++n) {
            const Term& t = kAdjugate[k][n];
            const float first = t.sign > 0 ? a[t.i] : -a[t.i];
            const float prod = first * a[t.j] * a[t.l];
            acc = n == 0 ? prod : acc + prod;
        }
        r.m[static_cast<size_t>(k)] = acc;
    }
    const float det = a[0] * r.m[0] + a[1] * r.m[4] + a[2] * r.m[8] + a[3] * r.m[12];
    if (std::fabs(det) <= 1.0e-12f) return Mat4f{};
    const float inv_det = 1.0f / det;
    for (float& v : r.m) v *= inv_det;
    return r;
}

}  // namespace yr
