// Variable / operator expression API of the explicit finite-volume path
// (include/fvm3d.hpp:36-530 of the reference): Variables, the algebra, the
// operators d1dx..d1dz (GradVar), d2dx..d2dz / lapl (Grad2Var), stabx..stab_tot
// (StabVar), the time operators d1t / d2t, DT with its update policies, and
// Equation::evaluate.
//
// Every Variable carries a shared expression node; copying or slicing a
// Variable (EquationTemplate stores the right-hand side by value) keeps the
// node, so subclasses never lose their meaning.  Evaluation runs on the
// MI355X: right-hand sides of equations flagged `true` are compiled into one
// fused kernel per equation (KernelBuilder, hipRTC); the others are evaluated
// operator by operator through the HIP library.  Within an equation the update
// is Jacobi (the new state goes to a separate buffer), equations run in order
// and boundary callbacks follow all equations, as in the reference's CPU path.
#pragma once

#include <functional>
#include <iostream>
#include <memory>
#include <optional>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "cuda_operators.hpp"
#include "decls.hpp"
#include "jit_config.h"
#include "mesh3d.hpp"

class Variable;
class DT;

using BoundaryFN = std::function<Eigen::Matrix<float, -1, 1>(Mesh3D* mesh, Eigen::Matrix<float, -1, 1>& arr, const DT* dt_)>;
using BoundaryFNCU = std::function<CudaDataMatrixD(Mesh3D* mesh, CudaDataMatrixD& arr, const DT* dt_)>;

using EquationTemplate = std::vector<std::tuple<Variable*, char, Variable, bool>>;

enum class TypeEnum { ConstDoublePointer, DoublePointer, ConstSizeTPointer, Boolean, SizeT, Double, Int };

namespace cfdarco {

enum class NodeKind { Leaf, Const, Array, Binary, Neg, Exp, Abs, Grad, Grad2, Stab, DtScalar, Pointer };

// Expression node.  Leaves point at the Variable that owns the state; the
// other kinds own their operands.
struct Node {
    NodeKind kind = NodeKind::Const;
    Variable* leaf = nullptr;           // Leaf
    float value = 0.0f;                 // Const
    CudaDataMatrixD array;              // Array (constant per-cell vector)
    char op = 0;                        // Binary: + - * /
    std::shared_ptr<Node> a, b;         // operands
    int mask = 7;                       // Grad / Grad2 / Stab axis mask (x=1, y=2, z=4)
    const DT* dt = nullptr;             // DtScalar
    const float* ptr = nullptr;         // Pointer
    Mesh3D* mesh = nullptr;
};

// per-op evaluation on the device (the reference's evaluate_cu tree walk)
CudaDataMatrixD evaluate_node(const Node& n);
size_t stencil_depth(const Node& n);

}  // namespace cfdarco

class Variable {
public:
    Variable();
    Variable(Mesh3D* mesh_, Eigen::Matrix<float, -1, 1>& initial_, BoundaryFN boundary_conditions_, std::string name_ = "");
    Variable(Mesh3D* mesh_, Eigen::Matrix<float, -1, 1>& initial_, BoundaryFN boundary_conditions_,
             BoundaryFNCU boundary_conditions_cu_, std::string name_ = "");
    Variable(Mesh3D* mesh_, float value);
    explicit Variable(Eigen::Matrix<float, -1, 1>& curr_);
    Variable(const Variable&);
    Variable& operator=(const Variable&);
    virtual ~Variable();

    [[nodiscard]] virtual std::shared_ptr<Variable> clone() const;

    void set_bound(const DT* dt_);
    void set_bound_cu(const DT* dt_);
    void add_history();

    virtual Eigen::Matrix<float, -1, 1> extract(Eigen::Matrix<float, -1, 1>& left_part, float dt);
    virtual CudaDataMatrixD extract_cu(CudaDataMatrixD& left_part, float dt);
    // host copy of the device evaluation
    virtual MatrixX4dRB evaluate();
    virtual CudaDataMatrixD evaluate_cu();

    void set_current(Eigen::Matrix<float, -1, 1>& current_);
    void set_current(CudaDataMatrixD& current_, bool copy_to_host);
    [[nodiscard]] std::vector<Eigen::Matrix<float, -1, 1>> get_history() const;

    virtual void solve(Variable* equation, DT* dt);

    Variable operator+(const Variable& obj_r) const;
    Variable operator-(const Variable& obj_r) const;
    Variable operator*(const Variable& obj_r) const;
    Variable operator/(const Variable& obj_r) const;
    Variable operator-() const;
    [[nodiscard]] Variable exp() const;

    virtual inline void set_dt(float) {}
    virtual inline void jit_post() {}

    std::string name;
    Mesh3D* mesh = nullptr;
    Eigen::Matrix<float, -1, 1> current;
    CudaDataMatrixD current_cu;
    BoundaryFN boundary_conditions;
    BoundaryFNCU boundary_conditions_cu;
    std::vector<Eigen::Matrix<float, -1, 1>> history{};
    size_t num_nodes = 0;
    bool has_boundary_conditions_cu = false;
    bool is_subvariable = false;
    bool is_constvar = false;
    bool is_basically_created = false;
    bool is_dt2 = false;

    std::shared_ptr<cfdarco::Node> node;

    // a Variable denoting an expression node (used by the algebra and operators)
    static Variable from_node(std::shared_ptr<cfdarco::Node> n, std::string name_);

protected:
    Variable(std::shared_ptr<cfdarco::Node> n, std::string name_);
};

Variable operator+(float obj_l, const Variable& obj_r);
Variable operator-(float obj_l, const Variable& obj_r);
Variable operator*(float obj_l, const Variable& obj_r);
Variable operator/(float obj_l, const Variable& obj_r);
Variable operator+(const Variable& obj_l, float obj_r);
Variable operator-(const Variable& obj_l, float obj_r);
Variable operator*(const Variable& obj_l, float obj_r);
Variable operator/(const Variable& obj_l, float obj_r);
Variable exp(const Variable& obj);
Variable abs(const Variable& obj);

class UpdatePolicies {
public:
    static float CourantFriedrichsLewy(float CFL, std::vector<Eigen::Matrix<float, -1, 1>>& space_vars, Mesh3D* mesh);
    static float CourantFriedrichsLewy3D(float CFL, std::vector<Eigen::Matrix<float, -1, 1>>& space_vars, Mesh3D* mesh);
    static float CourantFriedrichsLewy1D(float CFL, std::vector<Eigen::Matrix<float, -1, 1>>& space_vars, Mesh3D* mesh);
    static float CourantFriedrichsLewy3DCu(float CFL, std::vector<CudaDataMatrixD*>& space_vars, Mesh3D* mesh);
    static float CourantFriedrichsLewy1DCu(float CFL, std::vector<CudaDataMatrixD*>& space_vars, Mesh3D* mesh);
    static float CourantFriedrichsLewyCu(float CFL, std::vector<CudaDataMatrixD>& space_vars, Mesh3D* mesh);
    static inline float constant_dt(float CFL, std::vector<Eigen::Matrix<float, -1, 1>>&, Mesh3D*) { return CFL; }
    static inline float constant_dt_cu(float CFL, std::vector<CudaDataMatrixD*>&, Mesh3D*) { return CFL; }
};

class PointerVariable : public Variable {
public:
    PointerVariable(Mesh3D* mesh_, float* ptr);
    [[nodiscard]] std::shared_ptr<Variable> clone() const override;
    float* _ptr = nullptr;
};

class DT : public Variable {
public:
    using UpdateFn = std::function<float(float, std::vector<Eigen::Matrix<float, -1, 1>>&, Mesh3D* mesh)>;
    using UpdateFnCu = std::function<float(float, std::vector<CudaDataMatrixD*>&, Mesh3D* mesh)>;

    DT(Mesh3D* mesh_, UpdateFn update_fn_, float CFL_, std::vector<Variable*> space_vars_);
    DT(Mesh3D* mesh_, UpdateFn update_fn_, UpdateFnCu update_fn_cu_, float CFL_, std::vector<Variable*> space_vars_);
    DT(Mesh3D* mesh_, UpdateFn update_fn_, float value_) : DT(mesh_, std::move(update_fn_), value_, {}) {}
    DT(Mesh3D* mesh_, UpdateFn update_fn_, UpdateFnCu update_fn_cu_, float value_)
        : DT(mesh_, std::move(update_fn_), std::move(update_fn_cu_), value_, {}) {}

    // computes the step's dt at the start of the step (src/fvm3d.cpp:641-662)
    void update();
    [[nodiscard]] std::shared_ptr<Variable> clone() const override;

    UpdateFn update_fn;
    UpdateFnCu update_fn_cu;
    std::vector<Variable*> space_vars;
    float _dt = 0.0;
    float CFL = 0.0;
    bool has_update_fn_cu = false;
    size_t _current_time_step_int = 0;
    float _current_time_dbl = 0.0;
};

// d1t(var): explicit Euler, var <- dt * rhs + var (src/fvm3d.cpp:675-708)
class DtVar : public Variable {
public:
    DtVar(Variable* var_, int);
    Eigen::Matrix<float, -1, 1> extract(Eigen::Matrix<float, -1, 1>& left_part, float dt) override;
    CudaDataMatrixD extract_cu(CudaDataMatrixD& left_part, float dt) override;
    void solve(Variable* equation, DT* dt) override;
    [[nodiscard]] std::shared_ptr<Variable> clone() const override;
    inline void set_dt(float dt) override { dt_ = dt; }

    float dt_ = 0;
    std::shared_ptr<Variable> var;
    CudaDataMatrixD next_current_cu;
};

// d2t(var): second-order time derivative (src/fvm3d.cpp:710-765).  In fused
// (flag true) equations: two Euler steps, then leapfrog
// dt*dt*rhs + 2*cur - prev.  In operator-by-operator equations: leapfrog once
// the variable's history holds two entries (history is kept when -s/--store is
// given), Euler otherwise - both as in the reference.
class D2tVar : public Variable {
public:
    D2tVar(Variable* var_, int);
    Eigen::Matrix<float, -1, 1> extract(Eigen::Matrix<float, -1, 1>& left_part, float dt) override;
    CudaDataMatrixD extract_cu(CudaDataMatrixD& left_part, float dt) override;
    void solve(Variable* equation, DT* dt) override;
    [[nodiscard]] std::shared_ptr<Variable> clone() const override;
    inline void set_dt(float dt) override {
        dt_ = dt;
        second_stage = second_stage_next;
        second_stage_next = second_stage_next_next;
        second_stage_next_next = 1;
    }

    std::shared_ptr<Variable> var;
    float dt_ = 0;
    size_t second_stage = 0;
    size_t second_stage_next = 0;
    size_t second_stage_next_next = 0;
    CudaDataMatrixD pre_current_cu;
    CudaDataMatrixD next_current_cu;
};

// Green-Gauss gradient with linear face interpolation and outward normals,
// summed over the selected axes (src/fvm3d.cpp:768-829)
class GradVar : public Variable {
public:
    explicit GradVar(Variable* var_, bool clc_x_ = true, bool clc_y_ = true, bool clc_z_ = true);
    [[nodiscard]] std::shared_ptr<Variable> clone() const override;
    std::shared_ptr<Variable> var;
    bool clc_x, clc_y, clc_z;
};

// Laplacian sum_f alpha_f (phi_nb - phi) n_axis A_f / V (src/fvm3d.cpp:831-906)
class Grad2Var : public Variable {
public:
    explicit Grad2Var(Variable* var_, bool clc_x_ = true, bool clc_y_ = true, bool clc_z_ = true);
    [[nodiscard]] std::shared_ptr<Variable> clone() const override;
    std::shared_ptr<Variable> var;
    bool clc_x, clc_y, clc_z;
};

// upwind-jump stabilisation (src/fvm3d.cpp:909-942)
class StabVar : public Variable {
public:
    explicit StabVar(Variable* var_, bool clc_x_ = true, bool clc_y_ = true, bool clc_z_ = true);
    [[nodiscard]] std::shared_ptr<Variable> clone() const override;
    std::shared_ptr<Variable> var;
    bool clc_x, clc_y, clc_z;
};

// Fused per-equation kernel (the reference's OCCA KernelBuilder,
// src/fvm3d.cpp:1218-1394): the right-hand side tree is generated as HIP
// source, compiled once by hipRTC (cached), and launched every step.
// Sub-expressions a single pass cannot fuse (a stencil inside a stencil
// operand, and the two-stage stabilisation) are evaluated operator by
// operator before the launch and passed in as buffers.
class KernelBuilder {
public:
    explicit KernelBuilder(Variable* left_val_, Variable* expression_);
    KernelBuilder(KernelBuilder&&) noexcept;
    KernelBuilder& operator=(KernelBuilder&&) noexcept;
    ~KernelBuilder();
    // analysis and code generation only (no device): fills source()/kernel_name()
    void generate();
    // compiles (or takes from the cache) the kernel; `announce` prints the
    // reference's "Kernel build" line
    void build(bool announce = true);
    void run();
    [[nodiscard]] const std::string& source() const;
    [[nodiscard]] const std::string& kernel_name() const;
    // kernels computing stabilised pointwise operands before the launch
    [[nodiscard]] const std::vector<std::string>& operand_sources() const;
    [[nodiscard]] const std::vector<std::string>& operand_kernel_names() const;

    Variable* expression;
    Variable* left_val;
    CudaDataMatrixD ret;

private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
};

inline auto d1t(Variable& var) {
    auto varr = new DtVar(&var, 0);
    varr->name = "d1t(" + var.name + ")";
    return varr;
}
inline auto d1t(Variable&& var) {
    auto varr = new DtVar(&var, 0);
    varr->name = "d1t(" + var.name + ")";
    return varr;
}
inline auto d2t(Variable& var) {
    auto varr = new D2tVar(&var, 0);
    varr->name = "d2t(" + var.name + ")";
    return varr;
}
inline auto d2t(Variable&& var) {
    auto varr = new D2tVar(&var, 0);
    varr->name = "d2t(" + var.name + ")";
    return varr;
}
inline auto d1dx(Variable& var) { return GradVar(&var, true, false, false); }
inline auto d1dx(Variable&& var) { return GradVar(&var, true, false, false); }
inline auto d1dy(Variable& var) { return GradVar(&var, false, true, false); }
inline auto d1dy(Variable&& var) { return GradVar(&var, false, true, false); }
inline auto d1dz(Variable& var) { return GradVar(&var, false, false, true); }
inline auto d1dz(Variable&& var) { return GradVar(&var, false, false, true); }
inline auto d2dx(Variable& var) { return Grad2Var(&var, true, false, false); }
inline auto d2dx(Variable&& var) { return Grad2Var(&var, true, false, false); }
inline auto d2dy(Variable& var) { return Grad2Var(&var, false, true, false); }
inline auto d2dy(Variable&& var) { return Grad2Var(&var, false, true, false); }
inline auto d2dz(Variable& var) { return Grad2Var(&var, false, false, true); }
inline auto d2dz(Variable&& var) { return Grad2Var(&var, false, false, true); }
inline auto lapl(Variable& var) { return Grad2Var(&var, true, true, true); }
inline auto lapl(Variable&& var) { return Grad2Var(&var, true, true, true); }
inline auto stab_tot(Variable& var) { return StabVar(&var, true, true, true); }
inline auto stab_tot(Variable&& var) { return StabVar(&var, true, true, true); }
inline auto stabx(Variable& var) { return StabVar(&var, true, false, false); }
inline auto stabx(Variable&& var) { return StabVar(&var, true, false, false); }
inline auto staby(Variable& var) { return StabVar(&var, false, true, false); }
inline auto staby(Variable&& var) { return StabVar(&var, false, true, false); }
inline auto stabz(Variable& var) { return StabVar(&var, false, false, true); }
inline auto stabz(Variable&& var) { return StabVar(&var, false, false, true); }

class EqSolver {
public:
    // evaluate the whole right-hand side, then extract and assign (Jacobi)
    static void solve_dt(Variable* equation, Variable* time_var, Variable* set_var, DT* dt);
};

class Equation {
public:
    explicit Equation(size_t timesteps_);
    void evaluate(std::vector<Variable*>& all_vars, EquationTemplate& equation_system, DT* dt, bool visualize,
                  std::vector<Variable*> store_vars = {}) const;
    size_t timesteps;
};
