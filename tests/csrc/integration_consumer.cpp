// A consumer of the reference's C++ interface, written as a reference user writes one
// (samples/tests/quasi.cpp:63-97 and the classes it reaches), built against the MI355X
// library instead of the reference's sources.  Besides solving,
// it goes through the interfaces a consumer keeping parts of its own ADMM uses:
//   * wrapd::Elements, through a base-class pointer (src/Elements.hpp:36-157);
//   * System::get_A / get_b and mcl::optlib::LinearSolver{compute, factorize, solve}
//     (deps/mcloptlib/include/MCL/LinearSolver.hpp:279-298) on the assembled system;
//   * wrapd::GlobalProblem as an mcl::optlib::SplitProblem (value, gradient);
//   * create_tets_from_mesh's vertex_offset (src/TetElements.hpp:250-288).
// Prints one line of "key value" pairs; with -s 1 also writes objectives.txt to -out.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <memory>
#include <random>
#include <sys/stat.h>

#include "wrapd_host.hpp"
#include "wrapd_meshio.hpp"
#include "wrapd_optlib.hpp"

int main(int argc, char** argv) {
    try {
        wrapd::Settings settings;
        if (settings.parse_args(argc, argv)) return 0;
        const std::vector<int> pins = wrapd::meshio::load_handles(settings.m_io.handles());
        wrapd::meshio::TetMeshData rest, init;
        if (!wrapd::meshio::load_elenode(rest, settings.m_io.rest_mesh()) ||
            !wrapd::meshio::load_elenode(init, settings.m_io.init_mesh()))
            return 1;
        const int n_v = static_cast<int>(rest.verts.size() / 3);
        const int n_e = static_cast<int>(rest.tets.size() / 4);

        // binding::add_tetmesh (samples/utils/AddMeshes.hpp:61-132)
        auto solver = std::make_shared<wrapd::Solver>(settings);
        wrapd::math::MatX3 X(n_v), Xi(n_v);
        std::copy(rest.verts.begin(), rest.verts.end(), X.v.begin());
        std::copy(init.verts.begin(), init.verts.end(), Xi.v.begin());
        solver->m_system->add_nodes(X, Xi, X);
        std::vector<int> sorted(pins);
        std::sort(sorted.begin(), sorted.end());
        std::vector<int> free_list(n_v), fix_list(n_v, -1);
        for (int i = 0, j = 0, k = 0; i < n_v; ++i) {
            if (j < static_cast<int>(sorted.size()) && i == sorted[j]) { free_list[i] = -1; fix_list[i] = j++; }
            else { free_list[i] = k++; }
        }
        wrapd::create_tets_from_mesh<double, wrapd::TetElements>(solver->m_system->tet_elements(), rest.verts.data(),
                                                                  rest.tets.data(), n_e, settings, free_list, fix_list, 0);
        solver->set_pins(sorted);
        solver->initialize(settings);
        solver->solve();
        wrapd::AlgorithmData* ad = solver->algorithm_data();
        std::cout << std::setprecision(17) << "iters " << ad->m_iter;

        // the element set through the reference's base class
        std::shared_ptr<wrapd::Elements> el = solver->m_system->tet_elements();
        std::shared_ptr<wrapd::TetElements> te = solver->m_system->tet_elements();
        std::cout << " elements " << el->num_elements() << " energy " << el->potential_energy_x()
                  << " polar_pe " << el->polar_potential_energy() << " last_local_pe " << te->m_curr_pe
                  << " penalty " << el->penalty_energy(false)
                  << " primal_residual_sq " << el->primal_residual_sq(settings.m_rot_awareness);

        // the global objective as a SplitProblem, at the solution: value, gradient, and a
        // central difference of the value along a random direction
        std::shared_ptr<wrapd::AlgorithmData> ad_shared(solver->algorithm_data(), [](wrapd::AlgorithmData*) {});
        wrapd::GlobalProblem problem(settings, solver->m_system, ad_shared);
        mcl::optlib::SplitProblem<double, -1>& sp = problem;
        wrapd::math::MatX3 x;
        te->get_x_free(x);
        wrapd::math::MatX3 g;
        const double v = sp.gradient(x, g);
        std::mt19937_64 rng(7);
        std::normal_distribution<double> nd;
        wrapd::math::MatX3 d(x.rows()), xp(x.rows()), xm(x.rows());
        double gd = 0.0;
        double dn = 0.0;
        for (size_t i = 0; i < d.v.size(); ++i) { d.v[i] = nd(rng); dn += d.v[i] * d.v[i]; }
        const double h = 1e-6 * std::sqrt(static_cast<double>(x.rows()) / dn);
        for (size_t i = 0; i < d.v.size(); ++i) {
            gd += g.v[i] * d.v[i];
            xp.v[i] = x.v[i] + h * d.v[i];
            xm.v[i] = x.v[i] - h * d.v[i];
        }
        const double fd = (sp.value(xp) - sp.value(xm)) / (2.0 * h);
        sp.value(x);   // back at the solution
        std::cout << " global_value " << v << " grad_fd_rel_err " << std::fabs(fd - gd) / std::max(1e-300, std::fabs(gd));

        // System::get_A / get_b and the LinearSolver, A's pattern ordered without coordinates
        wrapd::math::SpMat A;
        solver->m_system->get_A(A);
        mcl::optlib::LinearSolver ls(A);
        const wrapd::math::MatX3 b = solver->m_system->get_b();
        wrapd::math::VectorXd b0(b.rows()), b1(b.rows());
        for (int i = 0; i < b.rows(); ++i) { b0[i] = b(i, 0); b1[i] = b(i, 1); }
        const wrapd::math::VectorXd y0 = ls.solve(b0);
        ls.factorize(A);   // refactor on the same pattern
        const wrapd::math::VectorXd y1 = ls.solve(b1);
        auto residual = [&](const wrapd::math::VectorXd& y, const wrapd::math::VectorXd& rhs) {
            // A holds the upper triangle: r = (U + U^T - D) y - rhs
            std::vector<double> r(rhs.size());
            for (size_t i = 0; i < rhs.size(); ++i) r[i] = -rhs[i];
            const auto& op = A.outerIndex();
            const auto& in = A.innerIndex();
            const auto& av = A.valuesVec();
            for (int i = 0; i < A.rows(); ++i)
                for (int p = op[i]; p < op[i + 1]; ++p) {
                    const int j = in[p];
                    r[i] += av[p] * y[j];
                    if (j != i) r[j] += av[p] * y[i];
                }
            double rn = 0.0, bn = 0.0;
            for (size_t i = 0; i < r.size(); ++i) { rn += r[i] * r[i]; bn += rhs[i] * rhs[i]; }
            return std::sqrt(rn / bn);
        };
        std::cout << " A_nnz_upper " << A.nonZeros() << " linsolve_info " << static_cast<int>(ls.info())
                  << " linsolve_residual " << std::max(residual(y0, b0), residual(y1, b1));

        // vertex_offset: the element set's indices are shifted, the geometry is not
        std::shared_ptr<wrapd::TetElements> shifted;
        wrapd::create_tets_from_mesh<double, wrapd::TetElements>(shifted, rest.verts.data(), rest.tets.data(), n_e,
                                                                  settings, free_list, fix_list, 7);
        const bool offset_ok = shifted->inds()[0].x[0] == rest.tets[0] + 7 && shifted->num_elements() == n_e;
        std::cout << " vertex_offset_ok " << (offset_ok ? 1 : 0) << std::endl;

        if (settings.m_io.should_save_data()) {
            const std::string out = settings.m_io.output_dir();
            mkdir(out.c_str(), 0755);
            std::ofstream f(out + "/objectives.txt");
            const auto& obj = ad->per_iter_objectives();
            for (size_t i = 0; i < obj.size(); ++i)
                f << std::scientific << std::setprecision(17) << static_cast<double>(i) << " " << obj[i] << "\n";
        }
        return 0;
    } catch (const std::exception& e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }
}
