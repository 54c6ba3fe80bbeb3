// cam_image.h -- host images with a camera pose, as the reference's
// CamImage<T> (src/inc/cam_image.h:8-57) over Image<T, Standard>
// (src/inc/image.h:7-408): dense row-major storage (pitch = width*sizeof(T)),
// reset/copy helpers, element access, and the pose R, t (X_cam = R X_w + t,
// or camera-to-world with PlaneSweep::setAlternativeRelativeMatrixMethod).
#pragma once

#include <cstring>
#include <vector>

#include "structs.h"

namespace psw {

template <typename T>
struct CamImage {
    Matrix3D R;
    Vector3D t;

    CamImage() = default;
    CamImage(size_t w, size_t h) { reset(w, h); }

    size_t width() const { return w_; }
    size_t height() const { return h_; }
    size_t area() const { return w_ * h_; }
    size_t pitch() const { return w_ * sizeof(T); }
    bool isValid() const { return !data_.empty(); }

    void reset(size_t w, size_t h) {
        w_ = w;
        h_ = h;
        data_.assign(w * h, T());
    }
    void free() {
        data_.clear();
        data_.shrink_to_fit();
        w_ = h_ = 0;
    }
    T* data() { return data_.data(); }
    const T* data() const { return data_.data(); }
    T* rowPtr(size_t y = 0) { return data_.data() + y * w_; }
    const T* rowPtr(size_t y = 0) const { return data_.data() + y * w_; }
    T& operator()(size_t x, size_t y) { return data_[y * w_ + x]; }
    const T& operator()(size_t x, size_t y) const { return data_[y * w_ + x]; }
    T& operator[](size_t i) { return data_[i]; }
    const T& operator[](size_t i) const { return data_[i]; }

    // Copy from a host buffer with the given row pitch in bytes.
    void copyFrom(const T* ptr, size_t pitch) {
        for (size_t y = 0; y < h_; ++y)
            std::memcpy(rowPtr(y), reinterpret_cast<const unsigned char*>(ptr) + y * pitch, w_ * sizeof(T));
    }
    void copyTo(T* ptr, size_t pitch) const {
        for (size_t y = 0; y < h_; ++y)
            std::memcpy(reinterpret_cast<unsigned char*>(ptr) + y * pitch, rowPtr(y), w_ * sizeof(T));
    }

private:
    size_t w_ = 0, h_ = 0;
    std::vector<T> data_;
};

typedef unsigned char uchar;

}  // namespace psw
