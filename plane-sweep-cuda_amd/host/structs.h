// structs.h -- the reference's host math types (src/inc/structs.h,
// src/inc/helper_structs.h) for callers of libpsw, in namespace psw.
//
// Matrix3D is three float3 rows; products are row-dot-column with
// left-to-right sums and the inverse is the cofactor formula, so host code
// built on these types (with -ffp-contract=off) computes what the
// reference's host code computes.
#pragma once

#include <cmath>
#include <cstddef>
#include <ostream>

namespace psw {

struct float3 {
    float x, y, z;
};
struct float2 {
    float x, y;
};
struct int3 {
    int x, y, z;
};
inline float3 make_float3(float x, float y, float z) { return float3{x, y, z}; }

inline float3 operator+(float3 a, float3 b) { return make_float3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline float3 operator-(float3 a, float3 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline float3 operator-(float3 a) { return make_float3(-a.x, -a.y, -a.z); }
inline float3 operator*(float3 a, float3 b) { return make_float3(a.x * b.x, a.y * b.y, a.z * b.z); }
inline float3 operator*(float3 a, float b) { return make_float3(a.x * b, a.y * b, a.z * b); }
inline float3 operator*(float b, float3 a) { return make_float3(b * a.x, b * a.y, b * a.z); }
inline float3 operator/(float3 a, float b) { return make_float3(a.x / b, a.y / b, a.z / b); }
inline float3 operator+(float3 a, float b) { return make_float3(a.x + b, a.y + b, a.z + b); }
inline float3 operator-(float3 a, float b) { return make_float3(a.x - b, a.y - b, a.z - b); }
inline float dot(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// Launch shape of the reference's launchers (CUDA's dim3).
struct dim3 {
    unsigned x, y, z;
    dim3(unsigned x_ = 1, unsigned y_ = 1, unsigned z_ = 1) : x(x_), y(y_), z(z_) {}
};

struct Rectangle3D {
    float3 a, b;
    Rectangle3D() : a(make_float3(0, 0, 0)), b(make_float3(0, 0, 0)) {}
    Rectangle3D(float3 x, float3 y) : a(x), b(y) {}
    float3 size() const { return b - a; }
    float3 center() const { return (b + a) / 2.f; }
};

struct Vector3D {
    float x, y, z;
    Vector3D(float x_ = 0, float y_ = 0, float z_ = 0) : x(x_), y(y_), z(z_) {}
    Vector3D(const float v[3]) : x(v[0]), y(v[1]), z(v[2]) {}
    Vector3D(const double v[3]) : x((float)v[0]), y((float)v[1]), z((float)v[2]) {}
    Vector3D(float3 f) : x(f.x), y(f.y), z(f.z) {}
    operator float3() const { return make_float3(x, y, z); }
    float& operator()(unsigned char e) { return e == 2 ? z : (e == 1 ? y : x); }
    const float& operator()(unsigned char e) const { return e == 2 ? z : (e == 1 ? y : x); }
    float& at(unsigned char e) { return (*this)(e); }
    const float& at(unsigned char e) const { return (*this)(e); }
};

struct Matrix3D {
    float3 r[3];  // rows
    Matrix3D() { r[0] = r[1] = r[2] = make_float3(0, 0, 0); }
    Matrix3D(float3 r1, float3 r2, float3 r3) { r[0] = r1; r[1] = r2; r[2] = r3; }
    Matrix3D(const float m[3][3]) {
        for (int i = 0; i < 3; ++i) r[i] = make_float3(m[i][0], m[i][1], m[i][2]);
    }
    Matrix3D(const double m[3][3]) {
        for (int i = 0; i < 3; ++i) r[i] = make_float3((float)m[i][0], (float)m[i][1], (float)m[i][2]);
    }
    Matrix3D(const float m[9]) {
        for (int i = 0; i < 3; ++i) r[i] = make_float3(m[3 * i], m[3 * i + 1], m[3 * i + 2]);
    }
    Matrix3D(float r11, float r12, float r13, float r21, float r22, float r23, float r31, float r32, float r33) {
        r[0] = make_float3(r11, r12, r13);
        r[1] = make_float3(r21, r22, r23);
        r[2] = make_float3(r31, r32, r33);
    }
    float3& row(unsigned char i) { return r[i]; }
    const float3& row(unsigned char i) const { return r[i]; }
    Matrix3D trans() const {
        return Matrix3D(make_float3(r[0].x, r[1].x, r[2].x), make_float3(r[0].y, r[1].y, r[2].y),
                        make_float3(r[0].z, r[1].z, r[2].z));
    }
    float det() const {
        return r[0].x * (r[1].y * r[2].z - r[1].z * r[2].y) - r[0].y * (r[1].x * r[2].z - r[1].z * r[2].x) +
               r[0].z * (r[1].x * r[2].y - r[1].y * r[2].x);
    }
    Matrix3D inv() const {
        const float d = det();
        const float3 r1 = make_float3(r[1].y * r[2].z - r[1].z * r[2].y, r[0].z * r[2].y - r[0].y * r[2].z,
                                      r[0].y * r[1].z - r[0].z * r[1].y) / d;
        const float3 r2 = make_float3(r[1].z * r[2].x - r[1].x * r[2].z, r[0].x * r[2].z - r[0].z * r[2].x,
                                      r[0].z * r[1].x - r[0].x * r[1].z) / d;
        const float3 r3 = make_float3(r[1].x * r[2].y - r[1].y * r[2].x, r[0].y * r[2].x - r[0].x * r[2].y,
                                      r[0].x * r[1].y - r[0].y * r[1].x) / d;
        return Matrix3D(r1, r2, r3);
    }
    void makeIdentity() {
        r[0] = make_float3(1, 0, 0);
        r[1] = make_float3(0, 1, 0);
        r[2] = make_float3(0, 0, 1);
    }
    static Matrix3D identityMatrix() {
        Matrix3D m;
        m.makeIdentity();
        return m;
    }
    float& operator()(unsigned char row, unsigned char col) {
        return col == 0 ? r[row].x : (col == 1 ? r[row].y : (col == 2 ? r[row].z : r[0].x));
    }
    const float& operator()(unsigned char row, unsigned char col) const {
        return col == 0 ? r[row].x : (col == 1 ? r[row].y : (col == 2 ? r[row].z : r[0].x));
    }
    float3& operator()(unsigned char row) { return r[row]; }
    const float3& operator()(unsigned char row) const { return r[row]; }
    // row-major copy for the C ABI
    void toArray(float out[9]) const {
        for (int i = 0; i < 3; ++i) {
            out[3 * i] = r[i].x;
            out[3 * i + 1] = r[i].y;
            out[3 * i + 2] = r[i].z;
        }
    }
    friend std::ostream& operator<<(std::ostream& o, const Matrix3D& m) {
        return o << "[" << m(0, 0) << " " << m(0, 1) << " " << m(0, 2) << "; " << m(1, 0) << " " << m(1, 1) << " "
                 << m(1, 2) << "; " << m(2, 0) << " " << m(2, 1) << " " << m(2, 2) << "]\n";
    }
};

// helper_structs.h:130-246 operators
inline float3 operator*(const Matrix3D& R, float3 v) { return make_float3(dot(R(0), v), dot(R(1), v), dot(R(2), v)); }
inline Matrix3D operator*(const Matrix3D& A, Matrix3D B) {
    B = B.trans();
    Matrix3D o;
    for (int i = 0; i < 3; ++i) o.r[i] = make_float3(dot(A(i), B(0)), dot(A(i), B(1)), dot(A(i), B(2)));
    return o;
}
inline Matrix3D operator*(const Matrix3D& A, float b) { return Matrix3D(A.r[0] * b, A.r[1] * b, A.r[2] * b); }
inline Matrix3D operator*(float a, const Matrix3D& B) { return Matrix3D(B.r[0] * a, B.r[1] * a, B.r[2] * a); }
inline Matrix3D operator/(const Matrix3D& A, float b) { return Matrix3D(A.r[0] / b, A.r[1] / b, A.r[2] / b); }
inline Matrix3D operator+(const Matrix3D& a, const Matrix3D& b) {
    return Matrix3D(a.r[0] + b.r[0], a.r[1] + b.r[1], a.r[2] + b.r[2]);
}
inline Matrix3D operator-(const Matrix3D& a, const Matrix3D& b) {
    return Matrix3D(a.r[0] - b.r[0], a.r[1] - b.r[1], a.r[2] - b.r[2]);
}
inline Matrix3D operator+(const Matrix3D& A, float b) { return Matrix3D(A.r[0] + b, A.r[1] + b, A.r[2] + b); }
inline Matrix3D operator-(const Matrix3D& A, float b) { return Matrix3D(A.r[0] - b, A.r[1] - b, A.r[2] - b); }
inline void operator*=(Matrix3D& A, const Matrix3D& B) { A = A * B; }
inline void operator*=(Matrix3D& A, float b) { A = A * b; }
inline void operator/=(Matrix3D& A, float b) { A = A / b; }
inline void operator+=(Matrix3D& A, const Matrix3D& b) { A = A + b; }
inline void operator-=(Matrix3D& A, const Matrix3D& b) { A = A - b; }

}  // namespace psw
