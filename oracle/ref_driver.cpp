// Test-infrastructure driver that links the reference WRAPD sources as they lie
// under the reference tree (never copied) and dumps golden vectors.
//
// It is compiled only by oracle/Makefile into oracle/_ref/ and is used to pin the
// oracle (oracle/wrapd_oracle.cpp); nothing in the product links or runs it.
//
// Modes (first argument):
//   solve  <reference quasi flags...> [-E youngs] [-out dir]
//          Same binding as samples/tests/quasi.cpp:44-97, then writes per-iteration
//          objectives, reweight flags, inner L-BFGS counts, |dAL| of both steps and the
//          final positions (%.17g) into <dir>.  -E overrides Young's modulus, which the
//          reference hard-codes at 1e7 (src/Lame.hpp:38), through the public Lame setters.
//   prox   reads "s1 s2 s3 wsq_over_vol mu lambda l1 l2 l3" per line from stdin and runs
//          the reference prox exactly as src/TetElements.cpp:302-318 does
//          (mcl::optlib::LBFGS<double,3>, BacktrackingCubic, 20 iterations); prints
//          "x1 x2 x3 iters" with %.17g.
//   svd    reads 9 numbers (row-major F) per line; prints sigma(3) U(9, row-major)
//          V(9, row-major) from wrapd::math::svd (src/Math.hpp:140-161).
//   local  <reference quasi flags...> [-E youngs] [-n steps]
//          Times the reference local step alone: the same binding as `solve`, then
//          TetElements set up at the initial positions (deformation cache, dual variables
//          and weights as ADMM::solve sets them, src/ADMM.cpp:59-84) and `steps` calls of
//          TetElements::polar_update (src/TetElements.cpp:256-356), each after a
//          deformation-cache refresh.  No solver is initialized, so no matrix is factored.
//          Prints "elements n steps k local_s t" (t: the polar_update calls only).
//   hess   reads "s1 s2 s3 mu lambda" per line; prints the largest eigenvalue of
//          NHProx::hessian_density (src/TetElements.hpp:109-144) as the candidate weight
//          code does (src/TetElements.cpp:414-416).
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include <iostream>
#include <fstream>
#include <sstream>

#include "AddMeshes.hpp"
#include "MeshIO.hpp"

static int run_local(int argc, char** argv);

static int run_solve(int argc, char** argv) {
    std::vector<char*> args;
    args.push_back(argv[0]);
    double youngs = -1.0;
    std::string outdir = ".";
    int steps = 0;
    for (int i = 2; i < argc; ++i) {
        if (std::strcmp(argv[i], "-E") == 0 && i + 1 < argc) {
            youngs = std::atof(argv[++i]);
        } else if (std::strcmp(argv[i], "-out") == 0 && i + 1 < argc) {
            outdir = argv[++i];
        } else if (std::strcmp(argv[i], "-n") == 0 && i + 1 < argc) {
            steps = std::atoi(argv[++i]);
        } else {
            args.push_back(argv[i]);
        }
    }
    wrapd::Settings settings;
    settings.m_io.should_save_data(false);
    settings.parse_args(static_cast<int>(args.size()), args.data());
    if (youngs > 0) {
        // ModelSettings keeps the Lame object private but hands out a const reference
        // to an object that is not itself const; its public setters set mu and lambda.
        double nu_mu = 0, nu_lam = 0;
        {
            // Recover nu from the preset the parser left (E=1e7): lambda/mu = 2nu/(1-2nu).
            const wrapd::Lame& l = settings.m_model_settings.elastic_lame();
            double r = l.lambda() / l.mu();
            double nu = r / (2.0 * (1.0 + r));
            nu_mu = youngs / (2.0 * (1.0 + nu));
            nu_lam = youngs * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
        }
        wrapd::Lame& lm = const_cast<wrapd::Lame&>(settings.m_model_settings.elastic_lame());
        lm.mu(nu_mu);
        lm.lambda(nu_lam);
    }

    std::vector<int> pins;
    {
        std::ifstream in(settings.m_io.handles().c_str());
        std::string line;
        while (std::getline(in, line)) {
            std::stringstream ss(line);
            int idx;
            while (ss >> idx) pins.push_back(idx);
        }
    }
    auto solver = std::make_shared<wrapd::Solver>(settings);
    auto rest = mcl::TetMesh::create();
    auto init = mcl::TetMesh::create();
    mcl::meshio::load_elenode(rest.get(), settings.m_io.rest_mesh());
    mcl::meshio::load_elenode(init.get(), settings.m_io.init_mesh());
    std::shared_ptr<mcl::TetMesh> soln;
    if (!settings.m_io.soln_mesh().empty()) {
        soln = mcl::TetMesh::create();
        mcl::meshio::load_elenode(soln.get(), settings.m_io.soln_mesh());
    } else {
        soln = rest;
    }
    binding::add_tetmesh(solver.get(), rest, init, soln, pins, settings);
    if (steps > 0) {
        // local-step timing (mode `local`): ADMM::solve's set-up of the element state at X_init
        std::shared_ptr<wrapd::TetElements>& te = solver->m_system->tet_elements();
        const wrapd::math::MatX3& Xi = solver->m_system->X_init();
        std::vector<int> sp(pins);
        std::sort(sp.begin(), sp.end());
        const int nv = static_cast<int>(Xi.rows());
        wrapd::math::MatX3 x_free(nv - static_cast<int>(sp.size()), 3), x_fix(static_cast<int>(sp.size()), 3);
        for (int i = 0, j = 0, k = 0; i < nv; ++i) {
            if (j < static_cast<int>(sp.size()) && i == sp[j]) x_fix.row(j++) = Xi.row(i);
            else x_free.row(k++) = Xi.row(i);
        }
        te->update_fix_cache(x_fix);
        te->update_defo_cache(x_free);
        te->initialize_dual_vars(Xi);
        te->midupdate_weights(Xi);
        double local_s = 0.0;
        for (int it = 0; it < steps; ++it) {
            te->update_defo_cache(x_free);
            const auto t0 = std::chrono::steady_clock::now();
            te->polar_update(settings.m_reweighting == wrapd::Settings::Reweighting::ENABLED);
            local_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        }
        std::printf("elements %d steps %d local_s %.6f\n", te->num_elements(), steps, local_s);
        return 0;
    }
    solver->initialize(settings);
    solver->solve();

    wrapd::AlgorithmData* ad = solver->algorithm_data();
    auto wr = [&](const std::string& name, auto const& vec) {
        std::ofstream f(outdir + "/" + name);
        f.precision(17);
        for (size_t i = 0; i < vec.size(); ++i) f << vec[i] << "\n";
    };
    wr("objectives.txt", ad->per_iter_objectives());
    wr("reweighted.txt", ad->per_iter_reweighted());
    std::vector<int> inner(ad->m_per_iter_inner_iters.begin(),
                           ad->m_per_iter_inner_iters.begin() + ad->m_iter);
    wr("inner_iters.txt", inner);
    wr("ls_dal.txt", ad->m_per_iter_ls_delta_auglag);
    wr("gs_dal.txt", ad->m_per_iter_gs_delta_auglag);
    wr("x_errors.txt", ad->per_iter_state_x_errors());
    {
        const wrapd::math::MatX3& X = solver->m_system->X();
        std::ofstream f(outdir + "/final_x.txt");
        f.precision(17);
        for (int i = 0; i < X.rows(); ++i) f << X(i, 0) << " " << X(i, 1) << " " << X(i, 2) << "\n";
    }
    std::printf("local_s %.6f global_s %.6f refactor_s %.6f iters %d\n",
                ad->m_runtime.local_ms / 1000., ad->m_runtime.global_ms / 1000.,
                ad->m_runtime.refactor_ms / 1000., ad->m_iter);
    return 0;
}

static int run_prox() {
    double s1, s2, s3, wov, mu, lam, l1, l2, l3;
    while (std::cin >> s1 >> s2 >> s3 >> wov >> mu >> lam >> l1 >> l2 >> l3) {
        wrapd::Lame lame(1.0, 0.25);
        lame.mu(mu);
        lame.lambda(lam);
        mcl::optlib::LBFGS<double, 3> solver;
        solver.m_settings.ls_method = mcl::optlib::LSMethod::BacktrackingCubic;
        solver.m_settings.max_iters = 20;
        wrapd::NHProx problem;
        problem.set_lame(lame);
        problem.set_x0(wrapd::math::Vec3(s1, s2, s3));
        problem.set_wsq_over_volume(wov);
        wrapd::math::Vec3 x(l1, l2, l3);
        int it = solver.minimize(problem, x);
        std::printf("%.17g %.17g %.17g %d\n", x[0], x[1], x[2], it);
    }
    return 0;
}

static int run_svd() {
    double a[9];
    while (std::cin >> a[0] >> a[1] >> a[2] >> a[3] >> a[4] >> a[5] >> a[6] >> a[7] >> a[8]) {
        wrapd::math::Mat3x3 F;
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) F(r, c) = a[3 * r + c];
        wrapd::math::Vec3 s;
        wrapd::math::Mat3x3 U, V;
        wrapd::math::svd(F, s, U, V, true);
        std::printf("%.17g %.17g %.17g", s[0], s[1], s[2]);
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) std::printf(" %.17g", U(r, c));
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) std::printf(" %.17g", V(r, c));
        std::printf("\n");
    }
    return 0;
}

static int run_hess() {
    double s1, s2, s3, mu, lam;
    while (std::cin >> s1 >> s2 >> s3 >> mu >> lam) {
        wrapd::Lame lame(1.0, 0.25);
        lame.mu(mu);
        lame.lambda(lam);
        wrapd::NHProx problem;
        problem.set_lame(lame);
        wrapd::math::Mat3x3 K;
        problem.hessian_density(wrapd::math::Vec3(s1, s2, s3), K);
        double kmax = K.eigenvalues().real().maxCoeff();
        std::printf("%.17g\n", kmax);
    }
    return 0;
}

static int run_local(int argc, char** argv) {
    bool has_n = false;
    for (int i = 2; i < argc; ++i) has_n = has_n || std::strcmp(argv[i], "-n") == 0;
    if (!has_n) {
        std::fprintf(stderr, "local: -n <steps> is required\n");
        return 2;
    }
    return run_solve(argc, argv);
}

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: ref_driver solve|prox|svd|hess ...\n");
        return 2;
    }
    std::string mode = argv[1];
    if (mode == "solve") return run_solve(argc, argv);
    if (mode == "local") return run_local(argc, argv);
    if (mode == "prox") return run_prox();
    if (mode == "svd") return run_svd();
    if (mode == "hess") return run_hess();
    std::fprintf(stderr, "unknown mode %s\n", mode.c_str());
    return 2;
}
