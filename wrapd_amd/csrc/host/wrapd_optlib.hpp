// The optimisation-library interfaces the reference's global step is written against
// (deps/mcloptlib/include/MCL/SplitProblem.hpp:43-51, LinearSolver.hpp:279-298), over the
// MI355X solver, and the reference's adapter from System to SplitProblem
// (src/GlobalStepBehavior.hpp:35-73).
//
// LinearSolver factors a caller-given sparse SPD matrix with the GPU supernodal Cholesky
// (the numeric factorization and the triangular solves run on the device; the symbolic
// analysis on the host).  Like the reference's SimplicialLDLT<..., Upper>, it reads the
// upper triangle of A.
#pragma once

#include <memory>
#include <stdexcept>
#include <vector>

#include "wrapd_host.hpp"

namespace mcl {
namespace optlib {

// deps/mcloptlib/include/MCL/SplitProblem.hpp:43-51 (x: n x 3 positions)
template <typename Scalar, int DIM>
class SplitProblem {
 public:
    typedef wrapd::math::MatX3 MatX3;
    virtual ~SplitProblem() {}
    virtual bool converged(const MatX3& x0, const MatX3& x1, const MatX3& grad) = 0;
    virtual Scalar value(const MatX3& x) = 0;
    virtual Scalar gradient(const MatX3& x, MatX3& grad) = 0;
    virtual void gradient(MatX3& grad) = 0;
};

// the values of Eigen::ComputationInfo that LinearSolver::info reports
enum ComputationInfo { Success = 0, NumericalIssue = 1, NoConvergence = 2, InvalidInput = 3 };

// deps/mcloptlib/include/MCL/LinearSolver.hpp:279-298
class LinearSolver {
 public:
    LinearSolver() {}
    explicit LinearSolver(const wrapd::math::SpMat& A) { compute(A); }
    ~LinearSolver();
    LinearSolver(const LinearSolver&) = delete;
    LinearSolver& operator=(const LinearSolver&) = delete;

    void compute(const wrapd::math::SpMat& A);     // analyzePattern + factorize
    void factorize(const wrapd::math::SpMat& A);   // new values, the pattern of the last compute
    ComputationInfo info() const { return m_info; }
    wrapd::math::VectorXd solve(const wrapd::math::VectorXd& rhs);

    // extensions: three right-hand sides at once, coordinates of the unknowns for a
    // geometric nested dissection (otherwise the matrix graph's level structures order it),
    // and the HIP device
    wrapd::math::MatX3 solve3(const wrapd::math::MatX3& rhs);
    void set_coordinates(const std::vector<double>& xyz) { m_xyz = xyz; }
    void set_device(int device) { m_device = device; }

 private:
    void upper_values(const wrapd::math::SpMat& A, std::vector<double>& vals) const;
    wrapd_ctx* m_ctx = nullptr;
    int m_device = 0;
    int m_n = 0;
    std::vector<int> m_rp, m_cols;      // the analysed upper triangle
    std::vector<double> m_xyz;
    ComputationInfo m_info = InvalidInput;
};

}  // namespace optlib
}  // namespace mcl

namespace wrapd {

// src/GlobalStepBehavior.hpp:35-73: the global step's objective as a SplitProblem
class GlobalProblem : public mcl::optlib::SplitProblem<double, -1> {
 public:
    GlobalProblem(const Settings& settings, std::shared_ptr<System> system, std::shared_ptr<AlgorithmData> algorithm_data)
        : m_settings(settings), m_system(system), m_algorithm_data(algorithm_data) {}
    bool converged(const math::MatX3& x0, const math::MatX3& x1, const math::MatX3& grad) override {
        (void)x0;
        (void)x1;
        (void)grad;
        throw std::runtime_error("Error: we are not using the 'converged' function in the global step");
    }
    double value(const math::MatX3& x_free) override { return m_system->global_obj_value(x_free); }
    double gradient(const math::MatX3& x_free, math::MatX3& grad) override {
        return m_system->global_obj_grad(x_free, grad, true);
    }
    // at the current positions (the device holds them)
    void gradient(math::MatX3& grad) override { m_system->global_obj_grad(math::MatX3(), grad, false); }

 private:
    const Settings m_settings;
    std::shared_ptr<System> m_system;
    std::shared_ptr<AlgorithmData> m_algorithm_data;
};

}  // namespace wrapd
