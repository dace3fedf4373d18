// LinearSolver over the GPU supernodal Cholesky (see wrapd_optlib.hpp).
#include "wrapd_optlib.hpp"

#include <string>

#include "host_error.hpp"

namespace mcl {
namespace optlib {

LinearSolver::~LinearSolver() {
    if (m_ctx) wrapd_destroy(m_ctx);
}

// A's upper triangle (j >= i) in the analysed pattern's order; the pattern must match
void LinearSolver::upper_values(const wrapd::math::SpMat& A, std::vector<double>& vals) const {
    const std::vector<int>& op = A.outerIndex();
    const std::vector<int>& in = A.innerIndex();
    const std::vector<double>& av = A.valuesVec();
    vals.clear();
    vals.reserve(m_cols.size());
    for (int i = 0; i < m_n; ++i) {
        int q = m_rp[i];
        for (int p = op[i]; p < op[i + 1]; ++p) {
            if (in[p] < i) continue;
            if (q >= m_rp[i + 1] || m_cols[q] != in[p])
                throw std::runtime_error("LinearSolver::factorize: A's pattern differs from the analysed one");
            vals.push_back(av[p]);
            ++q;
        }
        if (q != m_rp[i + 1]) throw std::runtime_error("LinearSolver::factorize: A's pattern differs from the analysed one");
    }
}

void LinearSolver::compute(const wrapd::math::SpMat& A) {
    if (A.rows() != A.cols() || A.rows() <= 0) {
        m_info = InvalidInput;
        return;
    }
    m_n = A.rows();
    // the upper triangle's pattern; a missing diagonal entry is an explicit zero
    const std::vector<int>& op = A.outerIndex();
    const std::vector<int>& in = A.innerIndex();
    m_rp.assign(1, 0);
    m_cols.clear();
    for (int i = 0; i < m_n; ++i) {
        bool diag = false;
        for (int p = op[i]; p < op[i + 1]; ++p) diag = diag || in[p] == i;
        if (!diag) {
            m_info = InvalidInput;   // SimplicialLDLT would meet a zero pivot
            return;
        }
        for (int p = op[i]; p < op[i + 1]; ++p)
            if (in[p] >= i) m_cols.push_back(in[p]);
        m_rp.push_back(static_cast<int>(m_cols.size()));
    }
    if (m_ctx) wrapd_destroy(m_ctx);
    m_ctx = nullptr;
    int st = wrapd_create(&m_ctx, m_device);
    if (st != WRAPD_OK) throw wrapd::DeviceError(st, "LinearSolver: no usable HIP device");
    const double* xyz = m_xyz.size() == 3ull * m_n ? m_xyz.data() : nullptr;
    st = wrapd_linsolve_analyze(m_ctx, m_n, m_rp.data(), m_cols.data(), xyz);
    if (st != WRAPD_OK) throw wrapd::DeviceError(st, std::string("LinearSolver::compute: ") + wrapd_last_error(m_ctx));
    factorize(A);
}

void LinearSolver::factorize(const wrapd::math::SpMat& A) {
    if (!m_ctx || A.rows() != m_n) {
        m_info = InvalidInput;
        return;
    }
    std::vector<double> vals;
    upper_values(A, vals);
    const int st = wrapd_linsolve_factorize(m_ctx, vals.data());
    if (st == WRAPD_ERR_NOT_SPD) {
        m_info = NumericalIssue;
        return;
    }
    if (st != WRAPD_OK) throw wrapd::DeviceError(st, std::string("LinearSolver::factorize: ") + wrapd_last_error(m_ctx));
    m_info = Success;
}

wrapd::math::MatX3 LinearSolver::solve3(const wrapd::math::MatX3& rhs) {
    if (m_info != Success) throw std::runtime_error("LinearSolver::solve: no factorization");
    if (rhs.rows() != m_n) throw std::runtime_error("LinearSolver::solve: right-hand side of the wrong size");
    wrapd::math::MatX3 x(m_n);
    const int st = wrapd_solve3(m_ctx, rhs.data(), x.data());
    if (st != WRAPD_OK) throw wrapd::DeviceError(st, std::string("LinearSolver::solve: ") + wrapd_last_error(m_ctx));
    return x;
}

wrapd::math::VectorXd LinearSolver::solve(const wrapd::math::VectorXd& rhs) {
    if (static_cast<int>(rhs.size()) != m_n) throw std::runtime_error("LinearSolver::solve: right-hand side of the wrong size");
    wrapd::math::MatX3 b(m_n);
    for (int i = 0; i < m_n; ++i) b(i, 0) = rhs[i];
    const wrapd::math::MatX3 x = solve3(b);
    wrapd::math::VectorXd out(m_n);
    for (int i = 0; i < m_n; ++i) out[i] = x(i, 0);
    return out;
}

}  // namespace optlib
}  // namespace mcl
