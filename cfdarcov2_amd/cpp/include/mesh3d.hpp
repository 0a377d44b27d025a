// Structured hex mesh of the operator API (include/mesh3d.hpp:118-239 of the
// reference).  Construction, numbering and geometry follow the reference
// (Mesh3D(x,y,z,dx,dy,dz), init_basic_internals, compute); storage is the
// library's: one compact node record per cell for drivers that iterate
// `_nodes`, per-cell volume and boundary masks, and the device mesh owned by the
// HIP library (cfd_mesh_t).  The reference's N x 6 face tables are built only
// on request (compute_tables), since no kernel reads them.
#pragma once

#include <array>
#include <cstddef>
#include <cstdint>
#include <memory>
#include <unordered_map>
#include <vector>

#include "abstract_mesh.hpp"
#include "cuda_data_matrix.hpp"
#include "decls.hpp"

class Vertex3D;
class Face3D;
class Quadrangle3D;
class Mesh3D;

// three coordinates with the reference's (i) accessor
struct Coords3 {
    float v[3] = {0, 0, 0};
    float& operator()(int i) { return v[i]; }
    float operator()(int i) const { return v[i]; }
};

class Vertex3D : AbstractVertex {
public:
    Vertex3D(float x, float y, float z, size_t id) : _id{id} {
        _coords.v[0] = x;
        _coords.v[1] = y;
        _coords.v[2] = z;
    }
    void compute() override {}
    [[nodiscard]] Eigen::Matrix<float, -1, 1> coordinates() const override;
    [[nodiscard]] float x() const { return _coords(0); }
    [[nodiscard]] float y() const { return _coords(1); }
    [[nodiscard]] float z() const { return _coords(2); }

    Mesh3D* _mesh = nullptr;
    size_t _id;
    Coords3 _coords;
};

class Face3D : AbstractFace {
public:
    Face3D(size_t v1, size_t v2, size_t v3, size_t v4, size_t id) : _id{id}, _vertexes_id{v1, v2, v3, v4} {}
    void compute() override {}
    [[nodiscard]] bool is_boundary() const override { return _nodes_id.size() == 1; }

    Mesh3D* _mesh = nullptr;
    size_t _id;
    bool _set = false;
    std::array<size_t, 4> _vertexes_id;
    std::vector<size_t> _nodes_id{};
    Coords3 _normal, _normal_alt, _center_coords;
    float _area = 0;
};

// One cell: id, lattice coordinates, centre and volume; face and vertex ids
// come from the mesh's closed-form numbering.
class Quadrangle3D : AbstractCell {
public:
    Quadrangle3D() = default;
    void compute() override {}
    [[nodiscard]] Eigen::Matrix<float, -1, 1> center_coords() const override;
    [[nodiscard]] bool is_boundary() const override;
    [[nodiscard]] bool is_boundary_x() const;
    [[nodiscard]] bool is_boundary_y() const;
    [[nodiscard]] bool is_boundary_z() const;
    [[nodiscard]] float x() const { return _center_coords(0); }
    [[nodiscard]] float y() const { return _center_coords(1); }
    [[nodiscard]] float z() const { return _center_coords(2); }
    [[nodiscard]] std::array<size_t, 6> faces_id() const;
    [[nodiscard]] std::array<size_t, 8> vertexes_id() const;
    [[nodiscard]] static size_t opposite_face_id(size_t face_id);

    Mesh3D* _mesh = nullptr;
    size_t _id = 0;
    Coords3 _center_coords;
    std::array<size_t, 3> _coord_idx = {0, 0, 0};
    float _volume = 0;
};

class Mesh3D : AbstractMesh {
public:
    Mesh3D(size_t x, size_t y, size_t z, float dx, float dy, float dz)
        : _num_nodes{x * y * z}, _x{x}, _y{y}, _z{z}, _lx{dx * static_cast<float>(x)},
          _ly{dy * static_cast<float>(y)}, _lz{dz * static_cast<float>(z)}, _dx{dx}, _dy{dy}, _dz{dz} {}
    ~Mesh3D() override;

    void compute() override;
    void init_basic_internals();

    [[nodiscard]] size_t square_node_coord_to_idx(size_t x, size_t y, size_t z) const;
    [[nodiscard]] std::array<size_t, 6> square_node_coord_to_face_idx(size_t x, size_t y, size_t z) const;
    [[nodiscard]] std::array<size_t, 8> square_node_coord_to_vertex_idx(size_t x, size_t y, size_t z) const;

    // Non-uniform meshes (no counterpart in the reference, which only builds
    // the lattice): offsets added to the vertex coordinates before compute(),
    // V x 3 in the reference vertex order, or seeded uniform offsets in
    // [-frac*d, frac*d] on interior vertices (SplitMix64, as the Python
    // workload generator).
    void set_vertex_offsets(std::vector<double> xyz);
    void perturb_interior_vertices(uint64_t seed, double frac);
    [[nodiscard]] const std::vector<double>& vertex_offsets() const { return offsets_; }

    // Builds the reference's N x 6 tables (_ids, normals, areas, alpha_d, ...).
    void compute_tables();

    // device mesh of the HIP library (after CudaMesh3D::host_to_device)
    [[nodiscard]] cfd_mesh_t device_mesh() const { return dev_; }

    size_t _num_nodes;
    size_t _x, _y, _z;
    float _lx, _ly, _lz;
    float _dx, _dy, _dz;

    std::vector<std::shared_ptr<Vertex3D>> _vertexes{};   // not materialised
    std::vector<std::shared_ptr<Face3D>> _faces{};        // not materialised
    std::vector<Quadrangle3D*> _nodes{};

    MatrixX6dRB _normal_x{}, _normal_y{}, _normal_z{};
    std::vector<MatrixX6dRB*> _normals_all{};
    MatrixX6dRB _normal_alt_x{}, _normal_alt_y{}, _normal_alt_z{};
    std::vector<MatrixX6dRB*> _normals_alt_all{};
    MatrixX6Idx _ids{};
    MatrixX6SignIdx _ids_bound_free{};
    MatrixX6dRB _alpha_d{};
    std::vector<MatrixX6dRB> _n_min_alpha_d{};
    MatrixX6dRB _len_node_center_to_face{};

    Eigen::Matrix<float, -1, 1> _volumes{};
    MatrixX6dRB _face_areas{};
    Eigen::Matrix<float, -1, 1> _node_is_boundary{};
    Eigen::Matrix<float, -1, 1> _node_is_boundary_reverse{};

    static constexpr int n_dims = 3;
    static constexpr int n_faces = 6;

protected:
    std::vector<Quadrangle3D> node_store_;
    std::vector<double> offsets_;
    cfd_mesh_t dev_ = nullptr;
};

class CudaMesh3D : public Mesh3D {
public:
    using Mesh3D::Mesh3D;

    CudaDataMatrixD _normal_x_cu{}, _normal_y_cu{}, _normal_z_cu{};
    CudaDataMatrixD _normal_alt_x_cu{}, _normal_alt_y_cu{}, _normal_alt_z_cu{};
    CudaDataMatrix<size_t> _ids_cu{};
    CudaDataMatrix<ptrdiff_t> _ids_bound_free_cu{};
    CudaDataMatrixD _alpha_d_cu{};
    std::array<CudaDataMatrixD, n_dims> _n_min_alpha_d_cu{};
    CudaDataMatrixD _len_node_center_to_face_cu{};
    CudaDataMatrixD _volumes_cu{};
    CudaDataMatrixD _face_areas_cu{};
    CudaDataMatrixD _node_is_boundary_cu{};
    CudaDataMatrixD _node_is_boundary_reverse_cu{};

    // Creates the device mesh (geometry held by the HIP library, constants on
    // uniform meshes) and uploads the per-cell arrays; the N x 6 tables are
    // uploaded only if compute_tables() ran (CudaMesh3D::host_to_device,
    // include/mesh3d.hpp:209-236 of the reference).
    void host_to_device();
};
