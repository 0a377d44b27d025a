// Scalar and table types of the cfdARCO operator API (include/decls.hpp:25-32
// of the reference): float storage, N x 6 column-major face tables.
#pragma once

#include <Eigen/Dense>
#include <unsupported/Eigen/CXX11/Tensor>

using MatrixX4dRB = Eigen::Matrix<float, -1, -1, Eigen::ColMajor>;
using MatrixX6dRB = Eigen::Matrix<float, -1, -1, Eigen::ColMajor>;
using MatrixX6Idx = Eigen::Matrix<size_t, -1, -1, Eigen::ColMajor>;
using MatrixX6SignIdx = Eigen::Matrix<ptrdiff_t, -1, -1, Eigen::ColMajor>;
using TensorX6dRB = Eigen::Tensor<float, 3, Eigen::ColMajor>;

template <typename T>
using MatrixX6T = Eigen::Matrix<T, -1, -1, Eigen::ColMajor>;
