// Abstract mesh entity interfaces (include/abstract_mesh.hpp:24-57 of the
// reference).
#pragma once

#include <Eigen/Core>

class AbstractVertex {
public:
    virtual ~AbstractVertex() = default;
    virtual void compute() = 0;
    [[nodiscard]] virtual Eigen::Matrix<float, -1, 1> coordinates() const = 0;
};

class AbstractFace {
public:
    virtual ~AbstractFace() = default;
    virtual void compute() = 0;
    [[nodiscard]] virtual bool is_boundary() const = 0;
};

class AbstractCell {
public:
    virtual ~AbstractCell() = default;
    virtual void compute() = 0;
    [[nodiscard]] virtual Eigen::Matrix<float, -1, 1> center_coords() const = 0;
    [[nodiscard]] virtual bool is_boundary() const = 0;
};

class AbstractMesh {
public:
    virtual ~AbstractMesh() = default;
    virtual void compute() = 0;
};
