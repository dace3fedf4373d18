// quasi: the reference command line (samples/tests/quasi.cpp:44-180) on the MI355X solver.
// Reads TetGen .node/.ele meshes and a handles file, solves, and writes the reference's
// logs (objectives, x_errors, reweighted, accumulated_time_s: "%.14e %.14e" lines,
// samples/utils/DataLog.hpp:55-70) and final_mesh.{node,ele,obj}
// (samples/utils/MeshIO.hpp:131-179, 318-362) into the output directory (-out).
#include <sys/stat.h>

#include <algorithm>
#include <array>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "host/wrapd_host.hpp"
#include "host/wrapd_meshio.hpp"

namespace {

using wrapd::meshio::TetMeshData;
using wrapd::meshio::load_elenode;
using wrapd::meshio::save_elenode;
using wrapd::meshio::boundary_faces;
using wrapd::meshio::save_obj;

template <typename V>
void write_log(const std::string& path, const V& v) {
    std::ofstream f(path);
    if (!f) throw std::runtime_error("Error: Unable to open file.");
    for (size_t i = 0; i < v.size(); ++i)
        f << std::scientific << std::setprecision(14) << static_cast<double>(i) << " " << std::scientific
          << std::setprecision(14) << static_cast<double>(v[i]) << std::endl;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        wrapd::Settings settings;
        if (settings.parse_args(argc, argv)) return 0;
        const std::vector<int> pins = wrapd::meshio::load_handles(settings.m_io.handles());
        TetMeshData rest, init, soln;
        if (!load_elenode(rest, settings.m_io.rest_mesh()) || !load_elenode(init, settings.m_io.init_mesh())) return 1;
        if (settings.m_io.soln_provided()) { if (!load_elenode(soln, settings.m_io.soln_mesh())) return 1; }
        else soln = rest;
        if (rest.verts.size() != init.verts.size())
            throw std::runtime_error("Error: Rest mesh and init mesh do not have the same number of vertices.");
        const int n_v = static_cast<int>(rest.verts.size() / 3);
        const int n_e = static_cast<int>(rest.tets.size() / 4);

        wrapd::Solver solver(settings);
        wrapd::math::MatX3 X(n_v), Xi(n_v), Xs(n_v);
        std::copy(rest.verts.begin(), rest.verts.end(), X.v.begin());
        std::copy(init.verts.begin(), init.verts.end(), Xi.v.begin());
        std::copy(soln.verts.begin(), soln.verts.end(), Xs.v.begin());
        solver.m_system->add_nodes(X, Xi, Xs);
        std::vector<int> sorted(pins);
        std::sort(sorted.begin(), sorted.end());
        std::vector<int> free_list(n_v), fix_list(n_v, -1);
        int j = 0, k = 0;
        for (int i = 0; i < n_v; ++i) {
            if (j < static_cast<int>(sorted.size()) && i == sorted[j]) { free_list[i] = -1; fix_list[i] = j++; }
            else { free_list[i] = k++; }
        }
        wrapd::create_tets_from_mesh<double, wrapd::TetElements>(solver.m_system->tet_elements(), rest.verts.data(),
                                                                  rest.tets.data(), n_e, settings, free_list, fix_list, 0);
        solver.set_pins(sorted);
        if (settings.m_io.verbose() > 0)
            std::cout << "Added mesh: \n\tvertices: " << n_v << "\n\ttets: " << n_e << "\n\tpins: " << sorted.size()
                      << std::endl;
        solver.initialize(settings);
        solver.solve();

        wrapd::AlgorithmData* ad = solver.algorithm_data();
        const double local_s = ad->m_runtime.local_ms / 1000., global_s = ad->m_runtime.global_ms / 1000.,
                     refactor_s = ad->m_runtime.refactor_ms / 1000.;
        if (settings.m_io.verbose() > 0) {
            std::cout << std::fixed << "Final accumulated time (s): " << (local_s + global_s + refactor_s) << std::endl
                      << "-- Local step (s): " << local_s << std::endl
                      << "-- Global step (s): " << global_s << std::endl
                      << "-- Refactor step (s): " << refactor_s << std::endl;
        }
        const std::string out = settings.m_io.output_dir();
        mkdir(out.c_str(), 0755);
        if (settings.m_io.should_save_data()) {
            write_log(out + "/objectives.txt", ad->per_iter_objectives());
            write_log(out + "/x_errors.txt", ad->per_iter_state_x_errors());
            write_log(out + "/reweighted.txt", ad->per_iter_reweighted());
            write_log(out + "/accumulated_time_s.txt", ad->per_iter_accumulated_time_s());
        }
        TetMeshData fin;
        fin.tets = init.tets;
        fin.verts = solver.m_system->X().v;
        save_elenode(fin, out + "/final_mesh");
        save_obj(fin.verts, boundary_faces(fin.tets), out + "/final_mesh.obj");
        return 0;
    } catch (const wrapd::DeviceError& e) {
        std::cerr << e.what() << std::endl;
        return 10 + e.status;
    } catch (const std::exception& e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }
}
