// Device vector of the operator API (include/cuda_data_matrix.hpp:35-190 of
// the reference): a shared_ptr-owned buffer from the library's stream-ordered
// pool, through a cache of freed blocks.  Copy construction deep-copies, copy assignment shares (as the
// reference's implicit assignment does), from_eigen / to_eigen copy to and
// from host vectors.  All transfers are ordered on the null stream.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <memory>
#include <stdexcept>
#include <vector>

#include "cfdarcho_main_3d.hpp"
#include "decls.hpp"

namespace cfdarco {
void* device_alloc(size_t bytes);
void device_free(void* p, size_t bytes);
void copy_h2d(void* dst, const void* src, size_t bytes);
void copy_d2h(void* dst, const void* src, size_t bytes);
void copy_d2d(void* dst, const void* src, size_t bytes);
void fill_device(void* dst, const void* value, size_t elem_bytes, size_t n);
}  // namespace cfdarco

template <typename T>
class CudaDeleter {
public:
    explicit CudaDeleter(size_t size) : _size{size} {}
    void operator()(void* p) const {
        if (_size > 0 && p != nullptr) cfdarco::device_free(p, _size * sizeof(T));
    }
    size_t _size;
};

template <typename T>
class CudaDataMatrix {
public:
    CudaDataMatrix() : _size(0), data(nullptr, CudaDeleter<T>{0}) {}

    explicit CudaDataMatrix(size_t size) : _size{size} {
        data = std::shared_ptr<T>(static_cast<T*>(size ? cfdarco::device_alloc(size * sizeof(T)) : nullptr),
                                  CudaDeleter<T>{size});
    }

    CudaDataMatrix(const CudaDataMatrix& oth) : CudaDataMatrix(oth._size) {
        if (oth.data != nullptr && _size) cfdarco::copy_d2d(data.get(), oth.data.get(), _size * sizeof(T));
    }
    CudaDataMatrix(CudaDataMatrix&&) noexcept = default;
    CudaDataMatrix& operator=(const CudaDataMatrix&) = default;
    CudaDataMatrix& operator=(CudaDataMatrix&&) noexcept = default;

    CudaDataMatrix(size_t size, T const_val) : CudaDataMatrix(size) {
        if (_size) cfdarco::fill_device(data.get(), &const_val, sizeof(T), _size);
    }

    Eigen::Vector<T, -1> to_eigen() {
        Eigen::Vector<T, -1> ret{_size};
        if (_size) cfdarco::copy_d2h(ret.data(), data.get(), _size * sizeof(T));
        return ret;
    }

    static CudaDataMatrix from_eigen(MatrixX6T<T> mtrx) {
        const auto n = static_cast<size_t>(mtrx.size());
        CudaDataMatrix ret{n};
        if (n) cfdarco::copy_h2d(ret.data.get(), mtrx.data(), n * sizeof(T));
        return ret;
    }

    size_t _size;
    std::shared_ptr<T> data;
};

using CudaDataMatrixD = CudaDataMatrix<float>;
