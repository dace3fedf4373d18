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

namespace {

struct TetMeshData {
    std::vector<double> verts;   // (n,3)
    std::vector<int> tets;       // (m,4)
};

// samples/utils/MeshIO.hpp:182-315: header count, "<idx> ..." lines, 1-based detection
bool load_elenode(TetMeshData& m, const std::string& file) {
    {
        std::ifstream f(file + ".ele");
        if (!f) { std::cerr << "\n**TetMesh Error: Could not load " << file << ".ele" << std::endl; return false; }
        std::string header;
        std::getline(f, header);
        int n = 0;
        std::stringstream(header) >> n;
        m.tets.assign(4ull * n, 0);
        std::vector<char> set(n, 0);
        bool one = false;
        for (int i = 0; i < n; ++i) {
            std::string line;
            std::getline(f, line);
            std::stringstream ss(line);
            long idx;
            int a[4];
            ss >> idx >> a[0] >> a[1] >> a[2] >> a[3];
            if (i == 0 && idx == 1) one = true;
            if (one) { idx -= 1; for (int& x : a) x -= 1; }
            if (idx < 0 || idx >= n) { std::cerr << "\n**TetMesh Error: Your indices are bad" << std::endl; return false; }
            for (int q = 0; q < 4; ++q) m.tets[4 * idx + q] = a[q];
            set[idx] = 1;
        }
        for (char s : set) if (!s) { std::cerr << "\n**TetMesh Error: Your indices are bad" << std::endl; return false; }
    }
    {
        std::ifstream f(file + ".node");
        if (!f) { std::cerr << "\n**TetMesh Error: Could not load " << file << ".node" << std::endl; return false; }
        std::string header;
        std::getline(f, header);
        int n = 0;
        std::stringstream(header) >> n;
        m.verts.assign(3ull * n, 0.0);
        std::vector<char> set(n, 0);
        bool one = false;
        for (int i = 0; i < n; ++i) {
            std::string line;
            std::getline(f, line);
            std::stringstream ss(line);
            long idx;
            double x, y, z;
            ss >> idx >> x >> y >> z;
            if (i == 0 && idx == 1) one = true;
            if (one) idx -= 1;
            if (idx < 0 || idx >= n) { std::cerr << "\n**TetMesh Error: Your indices are bad" << std::endl; return false; }
            m.verts[3 * idx] = x; m.verts[3 * idx + 1] = y; m.verts[3 * idx + 2] = z;
            set[idx] = 1;
        }
        for (char s : set) if (!s) { std::cerr << "\n**TetMesh Error: Your indices are bad" << std::endl; return false; }
    }
    if (m.verts.empty() || m.tets.empty()) throw std::runtime_error("\n**TetMesh Error: Problem loading files");
    return true;
}

void save_elenode(const TetMeshData& m, const std::string& file) {
    {
        std::ofstream f(file + ".ele");
        const size_t n = m.tets.size() / 4;
        f << n << " 4 0\n";
        for (size_t i = 0; i < n; ++i)
            f << "\t" << i << ' ' << m.tets[4 * i] << ' ' << m.tets[4 * i + 1] << ' ' << m.tets[4 * i + 2] << ' '
              << m.tets[4 * i + 3] << "\n";
        f << "# Generated by the MI355X WRAPD solver";
    }
    {
        std::ofstream f(file + ".node");
        const size_t n = m.verts.size() / 3;
        f << n << " 3 0 0\n";
        for (size_t i = 0; i < n; ++i)
            f << "\t" << i << ' ' << std::scientific << std::setprecision(15) << m.verts[3 * i] << ' ' << m.verts[3 * i + 1]
              << ' ' << m.verts[3 * i + 2] << "\n";
        f << "# Generated by the MI355X WRAPD solver";
    }
}

// TetMesh::need_faces (samples/utils/TetMesh.hpp:195-222): faces used by exactly one tet,
// with the reference's per-tet face orientations.
std::vector<std::array<int, 3>> boundary_faces(const std::vector<int>& tets) {
    std::map<std::array<int, 3>, std::pair<int, std::array<int, 3>>> cnt;
    const size_t n = tets.size() / 4;
    for (size_t t = 0; t < n; ++t) {
        const int p0 = tets[4 * t], p1 = tets[4 * t + 1], p2 = tets[4 * t + 2], p3 = tets[4 * t + 3];
        const std::array<int, 3> fs[4] = {{p0, p1, p3}, {p0, p2, p1}, {p0, p3, p2}, {p1, p2, p3}};
        for (const auto& f : fs) {
            std::array<int, 3> key = f;
            std::sort(key.begin(), key.end());
            auto it = cnt.find(key);
            if (it == cnt.end()) cnt[key] = {1, f};
            else it->second.first++;
        }
    }
    std::vector<std::array<int, 3>> out;
    for (const auto& kv : cnt) if (kv.second.first == 1) out.push_back(kv.second.second);
    return out;
}

void save_obj(const std::vector<double>& verts, const std::vector<std::array<int, 3>>& faces, const std::string& file) {
    std::ofstream f(file);
    f << "# Generated by the MI355X WRAPD solver";
    for (size_t i = 0; i < verts.size() / 3; ++i)
        f << "\nv " << std::scientific << std::setprecision(15) << verts[3 * i] << ' ' << verts[3 * i + 1] << ' '
          << verts[3 * i + 2];
    for (const auto& t : faces) f << "\nf " << t[0] + 1 << ' ' << t[1] + 1 << ' ' << t[2] + 1;
    f << "\n";
}

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
        // load_handles (samples/tests/quasi.cpp:146-159)
        std::vector<int> pins;
        {
            std::ifstream in(settings.m_io.handles());
            std::string line;
            while (std::getline(in, line)) {
                std::stringstream ss(line);
                int idx;
                while (ss >> idx) pins.push_back(idx);
            }
        }
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
