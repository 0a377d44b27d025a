/*
 * cfdarco.h — C-ABI of the MI355X explicit finite-volume library
 * (libcfdarco_hip.so).
 *
 * Every entry point returns int: CFD_OK (0) on success, a negative CFD_E* code
 * on failure, with a message available from cfd_last_error() on the calling
 * thread.  Nothing here calls exit() or lets a C++ exception cross the ABI.
 * Sizes are int64_t.  Device work is ordered on the caller's stream (a
 * hipStream_t passed as void*, NULL = the null stream); the library only
 * synchronises where a function returns a host value (noted per function).
 * Field arrays are device pointers of the mesh's dtype, N = nx*ny*nz cells in
 * the reference cell order id = nz*ny*x + nz*y + z.
 *
 * Each declaration names the reference interface it replaces
 * (int0vrt/cfdARCOV2 file:line).
 */
#ifndef CFDARCO_H
#define CFDARCO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CFD_OK 0
#define CFD_E_INVALID (-1)    /* bad argument (null pointer, size mismatch, bad enum) */
#define CFD_E_HIP (-2)        /* HIP runtime error */
#define CFD_E_NOMEM (-3)      /* pool or device allocation failed / pool limit reached */
#define CFD_E_UNSUPPORTED (-4)/* valid request this build does not implement */
#define CFD_E_COMPILE (-5)    /* hipRTC compilation failed (log in cfd_last_error) */
#define CFD_E_NONFINITE (-6)  /* a state check found NaN/Inf */

#define CFD_F32 0
#define CFD_F64 1

typedef struct cfd_ctx_s* cfd_ctx_t;
typedef struct cfd_mesh_s* cfd_mesh_t;
typedef struct cfd_program_s* cfd_program_t;
typedef struct cfd_batch_s* cfd_batch_t;

/* ---- lifetime --------------------------------------------------------------
 * Replaces CFDArcoGlobalInit::enable_cuda device selection and RMM pool set-up
 * (src/cfdarcho_main_3d.cpp:304-347) and gpuErrchk/exit error handling
 * (include/cuda_operators.hpp:29-36). */
int cfd_init(int device, cfd_ctx_t* out);
int cfd_destroy(cfd_ctx_t ctx);
const char* cfd_last_error(void);
int cfd_version(void);

/* ---- memory ----------------------------------------------------------------
 * Stream-ordered pool allocation; replaces CudaDataMatrix's RMM pool
 * allocate/deallocate (include/cuda_data_matrix.hpp:35-83) and
 * Allocator::cuda_mem_pool (include/pool_allocator.hpp:29-33).
 * cfd_pool_set_limit caps the bytes the pool may hand out (0 = no cap); an
 * allocation beyond it fails with CFD_E_NOMEM instead of reaching the device. */
int cfd_alloc(cfd_ctx_t ctx, int64_t bytes, void** dptr, void* stream);
int cfd_free(cfd_ctx_t ctx, void* dptr, void* stream);
int cfd_pool_set_limit(cfd_ctx_t ctx, int64_t bytes);
int cfd_pool_stats(cfd_ctx_t ctx, int64_t* bytes_in_use, int64_t* peak_bytes);

#define CFD_H2D 1
#define CFD_D2H 2
#define CFD_D2D 3
/* Replaces CudaDataMatrix::from_eigen / to_eigen / deep copy
 * (include/cuda_data_matrix.hpp:85-177).  D2H waits for the copy. */
int cfd_memcpy(cfd_ctx_t ctx, void* dst, const void* src, int64_t bytes, int kind, void* stream);
int cfd_stream_sync(void* stream);

/* ---- mesh ------------------------------------------------------------------
 * Replaces Mesh3D(x,y,z,dx,dy,dz) + init_basic_internals + compute
 * (include/mesh3d.hpp:120-124, src/mesh3d.cpp:151-385) and
 * CudaMesh3D::host_to_device (include/mesh3d.hpp:209-236).
 * vertex_offsets (nullable) holds V = (nx+1)(ny+1)(nz+1) xyz triples, in the
 * reference vertex order, added to the lattice coordinates x*dx before any
 * geometry is computed (the non-uniform-mesh hook).  The library computes the
 * reference geometry in the mesh dtype.  When every face of a family has the
 * same geometry the mesh is stored as constants (uniform kernels, no geometry
 * traffic); otherwise as face-ordered SoA arrays.  force_general = 1 forces the
 * SoA form. */
typedef struct cfd_hex_desc {
    int64_t nx, ny, nz;
    double dx, dy, dz;
    int dtype;
    int force_general;
    const double* vertex_offsets;
} cfd_hex_desc;

typedef struct cfd_mesh_info {
    int64_t n_cells, n_faces, n_vertices;
    int64_t nx, ny, nz;
    int dtype;
    int uniform;            /* 1 = constant-geometry kernels */
    int64_t device_bytes;   /* bytes of geometry held on the device */
} cfd_mesh_info;

int cfd_mesh_create_hex(cfd_ctx_t ctx, const cfd_hex_desc* desc, cfd_mesh_t* out);
int cfd_mesh_destroy(cfd_mesh_t mesh);
int cfd_mesh_get_info(cfd_mesh_t mesh, cfd_mesh_info* info);

/* ---- unstructured hex meshes ----------------------------------------------
 * A hex mesh given by its connectivity, in the layout of the reference's
 * mesh.json (include/io_operators_tmpl.hpp:161-212; the -m/--mesh input the
 * reference's readers take, include/io_operators_tmpl.hpp:215-247,
 * legacy/io_operators.cpp:170-205): vertex coordinates; per face its 4
 * vertices (the first three give its normal) and its cells (the second -1 on
 * the boundary); per cell its 6 faces in the local order z-, z+, y-, y+, x-,
 * x+ and its 8 vertices.  The library computes the reference geometry
 * (src/mesh3d.cpp:43-105,151-321) from it in the mesh dtype.  reorder = 1
 * renumbers the cells by reverse Cuthill-McKee and the faces by their lowest
 * new cell; fields passed to the operators are in that numbering
 * (cfd_mesh_permutation gives perm[new] = old; reorder = 0 keeps the
 * caller's).  cfd_grad, cfd_lapl, cfd_stab, cfd_heat_run, cfd_wave_run,
 * cfd_euler_run (one thread per cell; the face axis of the Rusanov geometry is
 * that of the face's local pair in its first cell) and cfd_mesh_table take
 * such a mesh; the other entry points return CFD_E_UNSUPPORTED for it.  Results equal
 * the reference semantics bit for bit in any numbering. */
typedef struct cfd_umesh_desc {
    int64_t n_cells, n_faces, n_vertices;
    const double* vertices;      /* V x 3 */
    const int64_t* face_verts;   /* F x 4 */
    const int64_t* face_cells;   /* F x 2 */
    const int64_t* cell_faces;   /* N x 6 */
    const int64_t* cell_verts;   /* N x 8 */
    int dtype;
    int reorder;
} cfd_umesh_desc;
int cfd_mesh_create_unstructured(cfd_ctx_t ctx, const cfd_umesh_desc* desc, cfd_mesh_t* out);
/* perm[new] = old for the cells (N long) and the faces (F long); either may
 * be NULL.  Lattice meshes give the identity. */
int cfd_mesh_permutation(cfd_mesh_t mesh, int64_t* cell_perm, int64_t* face_perm);
/* The reference-layout tables 0..15 of an unstructured mesh in the caller's
 * numbering, computed on the host without a device. */
int cfd_unstructured_table(const cfd_umesh_desc* desc, int table, void* host_out);

/* Reference-layout tables (N x 6 column-major, element (i,fc) at fc*N+i) copied
 * to host memory, for parity checks against Mesh3D's tables
 * (include/mesh3d.hpp:248-269).  ids/ids_bf are int64; the others are in the mesh
 * dtype; vol/is_bnd/ccx/ccy/ccz are N long. */
#define CFD_T_IDS 0
#define CFD_T_IDS_BF 1
#define CFD_T_NX 2
#define CFD_T_NY 3
#define CFD_T_NZ 4
#define CFD_T_NAX 5
#define CFD_T_NAY 6
#define CFD_T_NAZ 7
#define CFD_T_AREA 8
#define CFD_T_ALPHA 9
#define CFD_T_LEN 10
#define CFD_T_VOL 11
#define CFD_T_IS_BND 12
#define CFD_T_CCX 13
#define CFD_T_CCY 14
#define CFD_T_CCZ 15
/* Geometry of the Euler stepper (cfd_euler_run; no reference counterpart): per
 * (cell, fc) the +axis unit normal and the area of the face's vector area
 * 0.5 (p2 - p0) x (p3 - p1), N x 6 like the tables above, and the cell volume
 * built from those vector areas (N). */
#define CFD_T_FV_NX 16
#define CFD_T_FV_NY 17
#define CFD_T_FV_NZ 18
#define CFD_T_FV_AREA 19
#define CFD_T_FV_VOL 20
int cfd_mesh_table(cfd_mesh_t mesh, int table, void* host_out);
/* The same tables computed on the host straight from a descriptor, without a
 * device or context (host-only; used for CPU-side parity checks). */
int cfd_hex_table(const cfd_hex_desc* desc, int table, void* host_out);

/* ---- operators (out is caller-owned, N long, mesh dtype) --------------------
 * axis_mask bit 0/1/2 selects x/y/z; the selected components are summed.
 * GradVar (d1dx/d1dy/d1dz) evaluate_cu + eval_grad (src/fvm3d.cpp:825-829,
 * src/cuda_operators.cu:237-305).  wall_axis >= 0: reflective wall, the face
 * value on boundary faces normal to that axis is zero; -1: reference rule. */
int cfd_grad(cfd_mesh_t mesh, const void* phi, void* out, int axis_mask, int wall_axis, void* stream);
/* Grad2Var (lapl, d2dx..) evaluate_cu + eval_grad2 (src/fvm3d.cpp:902-906,
 * src/cuda_operators.cu:490-569). */
int cfd_lapl(cfd_mesh_t mesh, const void* phi, void* out, int axis_mask, void* stream);
/* StabVar (stab_tot, stabx..) evaluate_cu + eval_stab (src/fvm3d.cpp:938-942,
 * src/cuda_operators.cu:307-487).  Exactly uniform meshes run one pass; other
 * meshes use a 3N scratch from the pool for the first stage's gradients. */
int cfd_stab(cfd_mesh_t mesh, const void* phi, void* out, int axis_mask, void* stream);
/* cfd_stab without any device allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  General meshes keep the
 * first stage's gradients in a fixed 48 MiB workspace that the library holds
 * on each GPU until the process ends (3 N entries of the mesh dtype: at most
 * 2,097,152 cells in FP64 and 4,194,304 in FP32; larger general meshes get
 * CFD_E_INVALID before anything is launched).  Uniform meshes need no
 * workspace.  Every call on one GPU shares its workspace, so calls that
 * overlap in time on one GPU (other streams, other threads) race. */
int cfd_stab_noalloc(cfd_mesh_t mesh, const void* phi, void* out, int axis_mask, void* stream);

/* Elementwise kernels: add_mtrx/sub_mtrx/mul_mtrx/div_mtrx/neg_mtrx/
 * mul_mtrx(float)/div_const (src/cuda_operators.cu:27-136) plus the DtVar /
 * D2tVar updates (src/fvm3d.cpp:694-697,742-751). */
#define CFD_EW_ADD 0        /* out = a + b */
#define CFD_EW_SUB 1        /* out = a - b */
#define CFD_EW_MUL 2        /* out = a * b */
#define CFD_EW_DIV 3        /* out = a / b */
#define CFD_EW_NEG 4        /* out = -a */
#define CFD_EW_MULS 5       /* out = a * s */
#define CFD_EW_DIVS 6       /* out = a / s */
#define CFD_EW_SMUL 7       /* out = s * a */
#define CFD_EW_AXPY 8       /* out = s * a + b        (DtVar extract)   */
#define CFD_EW_LEAP 9       /* out = (s*s) * a + 2*b - c (D2tVar extract), c passed as aux */
#define CFD_EW_FILL 10      /* out = s */
#define CFD_EW_COPY 11      /* out = a */
#define CFD_EW_ADDS 12      /* out = a + s */
#define CFD_EW_SADD 13      /* out = s + a */
#define CFD_EW_SUBS 14      /* out = a - s */
#define CFD_EW_SSUB 15      /* out = s - a */
#define CFD_EW_SDIV 16      /* out = s / a */
/* exp and abs of a field (Variable::exp / exp / abs, src/fvm3d.cpp:472-494;
 * the reference evaluates them on the host only and throws on its device
 * path).  exp is the device libm's (within 1 ulp of the exact value in FP32
 * and FP64, so within 2 ulp of any other faithful exp such as the host's);
 * abs clears the sign bit and is exact. */
#define CFD_EW_EXP 17       /* out = exp(a) */
#define CFD_EW_ABS 18       /* out = |a| */
int cfd_ewise(cfd_ctx_t ctx, int op, int dtype, const void* a, const void* b, const void* aux, double s,
              void* out, int64_t n, void* stream);

/* CFL time step, UpdatePolicies::CourantFriedrichsLewy3D/1D(Cu) and cfl_cu
 * (src/fvm3d.cpp:497-567, src/cuda_operators.cu:138-235):
 * cfl * min_i dl * (1 / (sqrt((gamma p)/rho) + |u|)) over n cells, evaluated
 * in the dtype as the reference's CPU form.  v and w may be NULL (1D form).
 * Synchronises the stream to return the host value. */
int cfd_cfl_min(cfd_ctx_t ctx, int dtype, const void* u, const void* v, const void* w, const void* p,
                const void* rho, int64_t n, double dl, double cfl, double gamma, double* dt_out, void* stream);

/* Non-finite check: *bad = number of NaN/Inf entries in x (synchronises). */
int cfd_count_nonfinite(cfd_ctx_t ctx, int dtype, const void* x, int64_t n, int64_t* bad, void* stream);

/* ---- fused explicit steps ---------------------------------------------------
 * Heat: {d1t(T), '=', alpha * lapl(T)} (examples/main_3d_heat.cpp:89-91 with
 * the JIT kernel of src/fvm3d.cpp:1219-1394): T <- dt*(alpha*lapl(T)) + T.
 * alpha_field (nullable) is a per-cell coefficient.  Ping-pong between T_a and
 * T_b: after nsteps the state is in T_a when nsteps is even, T_b when odd.
 * On structured meshes a run of 32 steps or more replays a HIP graph of a
 * captured 32-step block (captured on the first such run, and again when
 * the fields, alpha, alpha_field, dt or the heat form change), then runs the
 * remaining steps; a stream the caller is capturing gets plain launches, and
 * CFDARCO_HEAT_GRAPH=0 in the environment turns the blocks off.  Results are
 * the same either way.  Calls on one mesh from several host threads at once
 * must be serialised by the caller. */
int cfd_heat_run(cfd_mesh_t mesh, void* T_a, void* T_b, double alpha, const void* alpha_field, double dt,
                 int64_t nsteps, void* stream);

/* First-order acoustic system, one fused launch per step, equations in order
 *   d1t(u) = -d1dx(p); d1t(v) = -d1dy(p); d1t(w) = -d1dz(p);
 *   d1t(p) = K * (d1dx(u) + d1dy(v) + d1dz(w))      (K = -c^2)
 * with reflective walls (normal velocity face value zero) and zero-gradient p.
 * state_a/state_b each hold {p, u, v, w}; after nsteps the state is in state_a
 * when nsteps is even, state_b when odd.  K_field (nullable) is per cell.
 * Replaces four KernelBuilder launches per step (src/fvm3d.cpp:1019-1032). */
int cfd_wave_run(cfd_mesh_t mesh, void* const state_a[4], void* const state_b[4], double K,
                 const void* K_field, double dt, int64_t nsteps, void* stream);

/* Compressible Euler with a Rusanov face flux, one fused launch per step (no
 * reference counterpart: the reference's Sod driver, examples/
 * main_3d_sod_shock.cpp:150-167, composes a central predictor and StabVar).
 * state_a/state_b each hold {rho, rho u, rho v, rho w, E}; after nsteps the
 * state is in state_a when nsteps is even, state_b when odd.  The time step
 * dt = cfl * min_i V_i / sum_f A_f (|u.n_f| + a_i) is recomputed on the device
 * from the state at the start of every step; dt_hist (device, nullable, nsteps
 * long, mesh dtype) receives each step's dt.  Walls are reflective. */
int cfd_euler_run(cfd_mesh_t mesh, void* const state_a[5], void* const state_b[5], double gamma, double cfl,
                  int64_t nsteps, void* dt_hist, void* stream);
/* Kernel shape of cfd_euler_run: variant 0 = one thread per cell (every face
 * flux evaluated by both of its cells), 1 = 2.5D march (default) with
 * tile_rows (2, 4 or 8; 2 or 4 in FP64) y-rows x 64 z-cells per workgroup and `chunk`
 * x-planes, each cell's primitives computed once per plane and shared through
 * LDS.  Both give bit-identical results. */
int cfd_mesh_set_euler_shape(cfd_mesh_t mesh, int variant, int tile_rows, int chunk);

/* ---- adjoint of the fused acoustic step (uniform meshes; no reference
 * counterpart) ---------------------------------------------------------------
 * cfd_wave_traces: nsteps fused acoustic steps as cfd_wave_run, recording p
 * after step t (1-based) at the sensor cells into traces[(t-1)*nsensors + j]
 * (device, mesh dtype, nullable) and, when checkpoints (device, nullable) is
 * given, copying the state {p,u,v,w} (4 N values each) before steps 0, every,
 * 2*every, ... into checkpoint k = t / every.  sensor_cells is a host array of
 * at most 64 distinct cell ids. */
int cfd_wave_traces(cfd_mesh_t mesh, void* const state_a[4], void* const state_b[4], double K,
                    const void* K_field, double dt, int64_t nsteps, const int64_t* sensor_cells, int nsensors,
                    void* traces, void* checkpoints, int64_t every, void* stream);
/* cfd_wave_adjoint: the transposed steps in reverse order.  adj_a holds the
 * adjoint of the final state {p,u,v,w} (normally zero) and trace_grad (device,
 * nsteps*nsensors, nullable) the gradient of the loss with respect to the
 * traces; after the call adj_a (nsteps even) or adj_b (nsteps odd) holds the
 * gradient with respect to the initial state.  grad_K (device, N, nullable)
 * accumulates dL/dK per cell and needs the checkpoints of cfd_wave_traces. */
int cfd_wave_adjoint(cfd_mesh_t mesh, void* const adj_a[4], void* const adj_b[4], double K,
                     const void* K_field, double dt, int64_t nsteps, const int64_t* sensor_cells, int nsensors,
                     const void* trace_grad, const void* checkpoints, int64_t every, void* grad_K, void* stream);

/* ---- adjoint of the heat step (uniform meshes; no reference counterpart) --
 * cfd_heat_traces: nsteps heat steps as cfd_heat_run, recording T after step t
 * (1-based) at the sensor cells into traces[(t-1)*nsensors + j] (device,
 * mesh dtype, nullable) and, when checkpoints (device, nullable) is given,
 * copying T (N values) before steps 0, every, 2*every, ... into checkpoint
 * k = t / every.  sensor_cells is a host array of at most 64 distinct ids. */
int cfd_heat_traces(cfd_mesh_t mesh, void* T_a, void* T_b, double alpha, const void* alpha_field, double dt,
                    int64_t nsteps, const int64_t* sensor_cells, int nsensors, void* traces, void* checkpoints,
                    int64_t every, void* stream);
/* cfd_heat_adjoint: the transposed steps H^T = I + dt L diag(k) in reverse
 * order (L is symmetric on a uniform mesh; k = alpha, or alpha_field).  adj_a
 * holds the adjoint of the final state (normally zero) and trace_grad
 * (device, nsteps*nsensors, nullable) the gradient of the loss with respect to
 * the traces; after the call adj_a (nsteps even) or adj_b (nsteps odd) holds
 * the gradient with respect to the initial state.  grad_alpha (device, N,
 * nullable) accumulates dL/dk per cell and needs the checkpoints of
 * cfd_heat_traces. */
int cfd_heat_adjoint(cfd_mesh_t mesh, void* adj_a, void* adj_b, double alpha, const void* alpha_field, double dt,
                     int64_t nsteps, const int64_t* sensor_cells, int nsensors, const void* trace_grad,
                     const void* checkpoints, int64_t every, void* grad_alpha, void* stream);

/* Kernel form of cfd_wave_traces / cfd_wave_adjoint (no reference
 * counterpart): form 0 = one thread per cell; form 1 = row form (uniform
 * meshes whose z extent is a multiple of 4 and at most 256, else
 * CFD_E_UNSUPPORTED) with `tile_rows` (4 or 8; 4 in FP64) output waves and
 * `chunk` x-planes per workgroup.  The row form's forward records the traces in the
 * fused step itself; its transposed step needs a constant K (runs with a K
 * field take form 0 for the transposed steps) and its K gradient keeps one
 * divergence field per recomputed step instead of three velocity fields.
 * Both forms give bit-identical traces and gradients.  Uniform meshes that allow it start on form 1 (4 waves, 8
 * planes). */
int cfd_mesh_set_adjoint_shape(cfd_mesh_t mesh, int form, int tile_rows, int chunk);

/* ---- batched ensemble (FP32 uniform meshes; no reference counterpart) -----
 * Many independent heat / acoustic cases advance together: one launch per
 * step covers every case's tiles, and cfd_batch_run(use_graph = 1) replays n
 * steps from a captured HIP graph with a single launch.  Each case's fields
 * equal a cfd_heat_run / cfd_wave_run of the same case bit for bit.  After n
 * steps a case's state is in state_a when n is even, state_b when odd; graph
 * runs need an even n.  The batch keeps pointers to the caller's meshes'
 * geometry constants (copied at creation) and state buffers. */
#define CFD_CASE_HEAT 0   /* d1t(T) = alpha * lapl(T), zero-gradient walls     */
#define CFD_CASE_WAVE 1   /* fused acoustic step with K = -c^2, reflective walls */
typedef struct cfd_batch_case {
    cfd_mesh_t mesh;      /* FP32 mesh, exactly uniform */
    int kind;             /* CFD_CASE_HEAT or CFD_CASE_WAVE */
    double coef;          /* heat: alpha; wave: K */
    double dt;
    void* state_a[4];     /* heat: [0] = T; wave: p, u, v, w */
    void* state_b[4];
} cfd_batch_case;
int cfd_batch_create(cfd_ctx_t ctx, const cfd_batch_case* cases, int ncases, cfd_batch_t* out);
/* Tile form: 0 = 2.5D march tiles (any extents; tile_rows 4, 8 or 16 y-rows x
 * 64 z-columns per workgroup), 1 = row form (256/Z z rows per wave, tile_rows
 * 4 or 8 output waves + 2 halo waves per workgroup; every case's z extent a
 * multiple of 4 and at most 256 and its fields 16-byte aligned, else
 * CFD_E_UNSUPPORTED), 2 = row form with two steps per launch (tile_rows = 8, 10
 * or 12 waves per workgroup; steps run in pairs of launches, the remaining
 * steps one per launch in form 1 with 4-row tiles).  chunk = x-planes per
 * tile. */
int cfd_batch_set_form(cfd_batch_t batch, int form);
int cfd_batch_set_shape(cfd_batch_t batch, int tile_rows, int chunk);
int cfd_batch_run(cfd_batch_t batch, int64_t nsteps, int use_graph, void* stream);
int cfd_batch_destroy(cfd_batch_t batch);

/* ---- runtime-compiled kernels --------------------------------------------
 * hipRTC compilation of HIP C++ source for gfx950 (-O3, no FMA contraction,
 * no packed FP32), cached by a hash of (source, kernel name, options) in
 * memory and on disk ($CFDARCO_JIT_CACHE, default ~/.cache/cfdarco/jit).
 * Replaces KernelBuilder::build / OCCA buildKernelFromString
 * (src/fvm3d.cpp:1219-1312).  Compile failures return CFD_E_COMPILE with the
 * compiler log in cfd_last_error(). */
int cfd_program_compile(cfd_ctx_t ctx, const char* src, const char* kernel_name, cfd_program_t* out);
/* Host-only compile (no device, no cache): code-object size in *code_bytes. */
int cfd_program_compile_only(const char* src, const char* kernel_name, int64_t* code_bytes);
/* Launch with `args` = array of pointers to the argument values, in the
 * kernel's parameter order.  Replaces KernelBuilder::run (src/fvm3d.cpp:1314-1394). */
int cfd_program_launch(cfd_program_t prog, const uint32_t grid[3], const uint32_t block[3], void** args,
                       void* stream);
int cfd_program_destroy(cfd_program_t prog);

/* Geometry as the kernels see it, for code generated at run time (the fused
 * kernels of the operator API).  Uniform meshes: per-axis constants (index 0/1/2
 * = x/y/z).  General meshes: device arrays per face family (0 = z-, 1 = y-,
 * 2 = x-faces) in the face order documented in cfdarcov2_amd/csrc/common.hpp,
 * holding the axis-direction normal, area and alpha_d, plus per-cell volume
 * and the six centre-to-face lengths (6 x N). */
typedef struct cfd_geometry {
    int uniform;
    int dtype;
    int64_t nx, ny, nz;
    double area[3], alpha[3], alpha_b[3], len[3], vol;
    const void* face_normal[3][3];
    const void* face_area[3];
    const void* face_alpha[3];
    const void* cell_vol;
    const void* cell_len;
} cfd_geometry;
int cfd_mesh_geometry(cfd_mesh_t mesh, cfd_geometry* out);

/* Kernel shape of the fused acoustic step on uniform meshes (no counterpart
 * in the reference, whose JIT kernels use fixed 64-thread blocks,
 * src/fvm3d.cpp:1306-1308):
 *   0  one thread per cell (the form for small meshes);
 *   1  2.5D marching kernel, `tile_rows` (4, 8, 16 or 32) y-rows of 64 z-cells
 *      per workgroup carried by 4 (8 for 32) thread rows; any z extent;
 *   7  row form (z extent a multiple of 4 and at most 256): one wave per z row
 *      with four cells per lane, `tile_rows` (4 or 8; 14 in FP32) output rows
 *      plus two halo rows per workgroup, non-temporal output stores (FP64
 *      runs with a K field take 4 rows where 8 are set: that build would
 *      spill registers);
 *   8  variant 7 with ordinary stores;
 *   9  variant 7 with non-temporal loads and ordinary stores;
 *   11 two steps per launch (`tile_rows` 4, 8 or 10 in FP32, 2 in FP64; steps
 *      run in pairs of launches, the remainder and K-field runs as variant 9);
 *   12 variant 11 with non-temporal stores.
 * `chunk` x-planes per workgroup; XCD-aware tile placement when bit 0 of
 * `xcd_remap` is set; for the row forms, bit 1 of `xcd_remap` makes every
 * other step walk its tiles in reverse.  The row forms return
 * CFD_E_UNSUPPORTED for other z extents; other variant numbers (2 to 6 and 10
 * were forms the autotuner never chose) return CFD_E_INVALID.
 * Uniform meshes whose z extent suits the row form start on variant 11 (FP32)
 * or 7 (FP64).
 * All variants give bit-identical results.  Set by the autotuner
 * (cfdarcov2_amd/autotune.py). */
int cfd_mesh_set_wave_shape(cfd_mesh_t mesh, int variant, int tile_rows, int chunk, int xcd_remap);

/* Kernel form of the heat step (cfd_heat_run) on exactly uniform meshes (no
 * counterpart in the reference): variant 0 = one thread per cell, one step per
 * launch; variant 1 = two steps per launch in row form (z extent a multiple of
 * 4 and at most 256, else CFD_E_UNSUPPORTED), `waves` (8, 12 or 16) waves per
 * workgroup and `chunk` x-planes per tile; runs go in pairs of launches, and
 * the remaining steps and runs with a diffusivity field take variant 0;
 * variant 2 = up to `chunk` (2 or 4) steps per launch in LDS boxes of `waves`
 * (4 or 8) x `waves` x min(Z, 64 - 2 chunk) output cells, each workgroup
 * computing its box widened by `chunk` cells for the steps in between; the
 * form for small meshes whose single step lasts about a launch.  All give bit-identical results.  Variant 1 pays on large meshes (it needs
 * enough tiles to fill the GPU): meshes of at least 2^23 cells that allow it
 * start on variant 1 (12 waves, 16 planes), others on variant 0;
 * cfdarcov2_amd/autotune.py:tune_heat times both. */
int cfd_mesh_set_heat_shape(cfd_mesh_t mesh, int variant, int waves, int chunk);

/* Autotuner (no counterpart in the reference, whose JIT kernels have one fixed
 * shape): times every kernel shape of the fused acoustic step (CFD_TUNE_WAVE,
 * the variants and tiles of cfd_mesh_set_wave_shape) or of the heat step
 * (CFD_TUNE_HEAT, cfd_mesh_set_heat_shape) on this mesh, `steps` steps each
 * (best of two runs, HIP events on the stream), on zeroed scratch fields from
 * the pool, and applies the fastest.  shape_out receives {variant,
 * tile_rows or waves, chunk, xcd_remap (0 for heat)}; us_per_step_out
 * (nullable) its time per step.  Synchronises the stream. */
#define CFD_TUNE_WAVE 0
#define CFD_TUNE_HEAT 1
int cfd_mesh_autotune(cfd_mesh_t mesh, int kind, int steps, int* shape_out, double* us_per_step_out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* CFDARCO_H */
