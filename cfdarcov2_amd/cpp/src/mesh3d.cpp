// Structured hex mesh of the operator API.  Numbering follows
// Mesh3D::square_node_coord_to_{idx,face_idx,vertex_idx} (src/mesh3d.cpp:387-416
// of the reference); geometry comes from the library's host builder
// (csrc/hexmesh.hpp), which restates Face3D/Quadrangle3D/Mesh3D::compute in
// the same operation order, and the device copy is the library's cfd_mesh_t.
#include <cstring>
#include <iostream>
#include <stdexcept>

#include "../../csrc/hexmesh.hpp"
#include "mesh3d.hpp"

namespace cfdarco {
void check(int rc, const char* what);
}

Eigen::Matrix<float, -1, 1> Vertex3D::coordinates() const {
    Eigen::Matrix<float, -1, 1> r{3};
    for (int i = 0; i < 3; ++i) r(i) = _coords(i);
    return r;
}

Eigen::Matrix<float, -1, 1> Quadrangle3D::center_coords() const {
    Eigen::Matrix<float, -1, 1> r{3};
    for (int i = 0; i < 3; ++i) r(i) = _center_coords(i);
    return r;
}

bool Quadrangle3D::is_boundary() const {
    return _coord_idx[0] == 0 || _coord_idx[0] + 1 == _mesh->_x || _coord_idx[1] == 0 ||
           _coord_idx[1] + 1 == _mesh->_y || _coord_idx[2] == 0 || _coord_idx[2] + 1 == _mesh->_z;
}
// boundary tests along one axis of this cell (the reference's versions test
// global faces 0-5 instead)
bool Quadrangle3D::is_boundary_x() const { return _coord_idx[0] == 0 || _coord_idx[0] + 1 == _mesh->_x; }
bool Quadrangle3D::is_boundary_y() const { return _coord_idx[1] == 0 || _coord_idx[1] + 1 == _mesh->_y; }
bool Quadrangle3D::is_boundary_z() const { return _coord_idx[2] == 0 || _coord_idx[2] + 1 == _mesh->_z; }

std::array<size_t, 6> Quadrangle3D::faces_id() const {
    return _mesh->square_node_coord_to_face_idx(_coord_idx[0], _coord_idx[1], _coord_idx[2]);
}
std::array<size_t, 8> Quadrangle3D::vertexes_id() const {
    return _mesh->square_node_coord_to_vertex_idx(_coord_idx[0], _coord_idx[1], _coord_idx[2]);
}

size_t Quadrangle3D::opposite_face_id(size_t face_id) {
    return face_id < 6 ? (face_id ^ 1u) : 0;
}

Mesh3D::~Mesh3D() {
    if (dev_ != nullptr && CFDArcoGlobalInit::context_or_null() != nullptr) cfd_mesh_destroy(dev_);
}

size_t Mesh3D::square_node_coord_to_idx(size_t x, size_t y, size_t z) const { return _z * _y * x + _z * y + z; }

std::array<size_t, 6> Mesh3D::square_node_coord_to_face_idx(size_t x, size_t y, size_t z) const {
    const size_t nzf = _x * _y * (_z + 1), nyf = _x * (_y + 1) * _z;
    return {(_z + 1) * _y * x + (_z + 1) * y + z,
            (_z + 1) * _y * x + (_z + 1) * y + (z + 1),
            nzf + (_x * (_y + 1) * z + (_y + 1) * x + y),
            nzf + (_x * (_y + 1) * z + (_y + 1) * x + (y + 1)),
            nzf + nyf + ((_x + 1) * _z * y + (_x + 1) * z + x),
            nzf + nyf + ((_x + 1) * _z * y + (_x + 1) * z + (x + 1))};
}

std::array<size_t, 8> Mesh3D::square_node_coord_to_vertex_idx(size_t x, size_t y, size_t z) const {
    const size_t sx = (_z + 1) * (_y + 1), sy = _z + 1;
    std::array<size_t, 8> r{};
    int k = 0;
    for (size_t a = 0; a < 2; ++a)
        for (size_t b = 0; b < 2; ++b)
            for (size_t c = 0; c < 2; ++c) r[k++] = (x + a) * sx + (y + b) * sy + (z + c);
    return r;
}

void Mesh3D::init_basic_internals() {
    node_store_.assign(_num_nodes, Quadrangle3D{});
    _nodes.resize(_num_nodes);
    for (size_t x = 0; x < _x; ++x)
        for (size_t y = 0; y < _y; ++y)
            for (size_t z = 0; z < _z; ++z) {
                const size_t id = square_node_coord_to_idx(x, y, z);
                Quadrangle3D& q = node_store_[id];
                q._mesh = this;
                q._id = id;
                q._coord_idx = {x, y, z};
                _nodes[id] = &q;
            }
}

void Mesh3D::set_vertex_offsets(std::vector<double> xyz) {
    if (xyz.size() != 3 * (_x + 1) * (_y + 1) * (_z + 1))
        throw std::invalid_argument("vertex offsets must hold 3 values per vertex");
    offsets_ = std::move(xyz);
}

void Mesh3D::perturb_interior_vertices(uint64_t seed, double frac) {
    // SplitMix64 (Steele, Lea & Flood 2014); doubles as (x >> 11) * 2^-53
    uint64_t state = seed;
    auto next = [&state]() {
        state += 0x9E3779B97F4A7C15ull;
        uint64_t z = state;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    const size_t V = (_x + 1) * (_y + 1) * (_z + 1);
    std::vector<double> off(3 * V, 0.0);
    const double d[3] = {double(_dx), double(_dy), double(_dz)};
    size_t v = 0;
    for (size_t x = 0; x <= _x; ++x)
        for (size_t y = 0; y <= _y; ++y)
            for (size_t z = 0; z <= _z; ++z, ++v) {
                const bool interior = x > 0 && x < _x && y > 0 && y < _y && z > 0 && z < _z;
                for (int k = 0; k < 3; ++k) {
                    const double u = double(next() >> 11) * (1.0 / 9007199254740992.0);
                    if (interior) off[3 * v + k] = (2.0 * u - 1.0) * frac * d[k];
                }
            }
    offsets_ = std::move(off);
}

void Mesh3D::compute() {
    if (node_store_.size() != _num_nodes) init_basic_internals();
    cfd::HexGeometry<float> g;
    cfd::build_hex_geometry<float>(g, int64_t(_x), int64_t(_y), int64_t(_z), _dx, _dy, _dz,
                                   offsets_.empty() ? nullptr : offsets_.data());
    std::cout << "NUM NODES = " << _num_nodes << std::endl;
    _volumes = Eigen::Matrix<float, -1, 1>{_num_nodes};
    _node_is_boundary = Eigen::Matrix<float, -1, 1>{_num_nodes};
    _node_is_boundary_reverse = Eigen::Matrix<float, -1, 1>{_num_nodes};
    for (size_t i = 0; i < _num_nodes; ++i) {
        Quadrangle3D& q = node_store_[i];
        q._center_coords(0) = g.ccx[i];
        q._center_coords(1) = g.ccy[i];
        q._center_coords(2) = g.ccz[i];
        q._volume = g.vol[i];
        _volumes(i) = g.vol[i];
        const bool b = q.is_boundary();
        _node_is_boundary(i) = b ? 1.f : 0.f;
        _node_is_boundary_reverse(i) = b ? 0.f : 1.f;
    }
    if (std::getenv("CFDARCO_MESH_TABLES")) compute_tables();
    std::cout << "Done mesh compute" << std::endl;
}

void Mesh3D::compute_tables() {
    cfd_hex_desc d{int64_t(_x), int64_t(_y), int64_t(_z), double(_dx), double(_dy), double(_dz), CFD_F32, 0,
                   offsets_.empty() ? nullptr : offsets_.data()};
    const size_t n6 = 6 * _num_nodes;
    auto ftab = [&](int which, MatrixX6dRB& m) {
        m = MatrixX6dRB{_num_nodes, 6};
        cfdarco::check(cfd_hex_table(&d, which, m.data()), "mesh table");
    };
    ftab(CFD_T_NX, _normal_x);
    ftab(CFD_T_NY, _normal_y);
    ftab(CFD_T_NZ, _normal_z);
    ftab(CFD_T_NAX, _normal_alt_x);
    ftab(CFD_T_NAY, _normal_alt_y);
    ftab(CFD_T_NAZ, _normal_alt_z);
    ftab(CFD_T_AREA, _face_areas);
    ftab(CFD_T_ALPHA, _alpha_d);
    ftab(CFD_T_LEN, _len_node_center_to_face);
    std::vector<int64_t> ids(n6);
    cfdarco::check(cfd_hex_table(&d, CFD_T_IDS, ids.data()), "mesh table");
    _ids = MatrixX6Idx{_num_nodes, 6};
    for (size_t k = 0; k < n6; ++k) _ids.data()[k] = size_t(ids[k]);
    cfdarco::check(cfd_hex_table(&d, CFD_T_IDS_BF, ids.data()), "mesh table");
    _ids_bound_free = MatrixX6SignIdx{_num_nodes, 6};
    for (size_t k = 0; k < n6; ++k) _ids_bound_free.data()[k] = ptrdiff_t(ids[k]);
    _normals_all = {&_normal_x, &_normal_y, &_normal_z};
    _normals_alt_all = {&_normal_alt_x, &_normal_alt_y, &_normal_alt_z};
}

void CudaMesh3D::host_to_device() {
    cfd_ctx_t ctx = CFDArcoGlobalInit::context();
    if (dev_ != nullptr) {
        cfd_mesh_destroy(dev_);
        dev_ = nullptr;
    }
    cfd_hex_desc d{int64_t(_x), int64_t(_y), int64_t(_z), double(_dx), double(_dy), double(_dz), CFD_F32, 0,
                   offsets_.empty() ? nullptr : offsets_.data()};
    cfdarco::check(cfd_mesh_create_hex(ctx, &d, &dev_), "cfd_mesh_create_hex");
    _volumes_cu = CudaDataMatrixD::from_eigen(_volumes);
    _node_is_boundary_cu = CudaDataMatrixD::from_eigen(_node_is_boundary);
    _node_is_boundary_reverse_cu = CudaDataMatrixD::from_eigen(_node_is_boundary_reverse);
    if (_ids.size() > 0) {
        _normal_x_cu = CudaDataMatrixD::from_eigen(_normal_x);
        _normal_y_cu = CudaDataMatrixD::from_eigen(_normal_y);
        _normal_z_cu = CudaDataMatrixD::from_eigen(_normal_z);
        _normal_alt_x_cu = CudaDataMatrixD::from_eigen(_normal_alt_x);
        _normal_alt_y_cu = CudaDataMatrixD::from_eigen(_normal_alt_y);
        _normal_alt_z_cu = CudaDataMatrixD::from_eigen(_normal_alt_z);
        _ids_cu = CudaDataMatrix<size_t>::from_eigen(_ids);
        _ids_bound_free_cu = CudaDataMatrix<ptrdiff_t>::from_eigen(_ids_bound_free);
        _alpha_d_cu = CudaDataMatrixD::from_eigen(_alpha_d);
        _len_node_center_to_face_cu = CudaDataMatrixD::from_eigen(_len_node_center_to_face);
        _face_areas_cu = CudaDataMatrixD::from_eigen(_face_areas);
    }
}
