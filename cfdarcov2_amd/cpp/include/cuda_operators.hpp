// Device operators of the operator API (include/cuda_operators.hpp:38-96 of
// the reference).  Each returns a freshly allocated result, as the reference
// does; the arithmetic runs in the HIP library (cfd_ewise, cfd_grad,
// cfd_lapl, cfd_stab, cfd_cfl_min) on the null stream.  Failures throw
// std::runtime_error with the library's message instead of calling exit().
#pragma once

#include "cuda_data_matrix.hpp"
#include "decls.hpp"
#include "mesh3d.hpp"

CudaDataMatrixD add_mtrx(const CudaDataMatrixD& a, const CudaDataMatrixD& b);
CudaDataMatrixD sub_mtrx(const CudaDataMatrixD& a, const CudaDataMatrixD& b);
CudaDataMatrixD mul_mtrx(const CudaDataMatrixD& a, const CudaDataMatrixD& b);
CudaDataMatrixD mul_mtrx(const CudaDataMatrixD& a, float b);
CudaDataMatrixD div_mtrx(const CudaDataMatrixD& a, const CudaDataMatrixD& b);
CudaDataMatrixD neg_mtrx(const CudaDataMatrixD& a);
CudaDataMatrixD div_const(const CudaDataMatrixD& a, float b);
// exp and abs of a field: no device form in the reference (its exp/abs run
// on the host only, src/fvm3d.cpp:472-494)
CudaDataMatrixD exp_mtrx(const CudaDataMatrixD& a);
CudaDataMatrixD abs_mtrx(const CudaDataMatrixD& a);

CudaDataMatrixD eval_grad2(CudaMesh3D* mesh, const CudaDataMatrixD& a, bool clc_x, bool clc_y, bool clc_z);
CudaDataMatrixD eval_grad(CudaMesh3D* mesh, const CudaDataMatrixD& a, bool clc_x, bool clc_y, bool clc_z);
CudaDataMatrixD eval_stab(CudaMesh3D* mesh, const CudaDataMatrixD& a, bool clc_x, bool clc_y, bool clc_z);

// min over cells of dl / (sqrt(gamma p / rho) + |u|) (3-velocity and 1D forms)
float cfl_cu(float dl, float gamma, const CudaDataMatrixD& p, const CudaDataMatrixD& rho, const CudaDataMatrixD& u,
             const CudaDataMatrixD& v, const CudaDataMatrixD& w);
float cfl_cu(float dl, float gamma, const CudaDataMatrixD& p, const CudaDataMatrixD& rho, const CudaDataMatrixD& u);

inline CudaDataMatrixD operator+(const CudaDataMatrixD& l, const CudaDataMatrixD& r) { return add_mtrx(l, r); }
inline CudaDataMatrixD operator-(const CudaDataMatrixD& l, const CudaDataMatrixD& r) { return sub_mtrx(l, r); }
inline CudaDataMatrixD operator*(const CudaDataMatrixD& l, const CudaDataMatrixD& r) { return mul_mtrx(l, r); }
inline CudaDataMatrixD operator/(const CudaDataMatrixD& l, const CudaDataMatrixD& r) { return div_mtrx(l, r); }
inline CudaDataMatrixD operator-(const CudaDataMatrixD& l) { return neg_mtrx(l); }

// waits for all device work (hipDeviceSynchronize), reporting errors by exception
void sync_device();
