// TetGen .node/.ele meshes, handles files and .obj surfaces, as the reference's sample
// utilities read and write them (samples/utils/MeshIO.hpp:131-362, TetMesh.hpp:195-222,
// samples/tests/quasi.cpp:146-159): header counts, "<idx> ..." lines, 0- or 1-based indices
// detected from the first line.
#pragma once

#include <algorithm>
#include <array>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace wrapd {
namespace meshio {

struct TetMeshData {
    std::vector<double> verts;   // (n,3)
    std::vector<int> tets;       // (m,4)
};

// samples/utils/MeshIO.hpp:182-315: header count, "<idx> ..." lines, 1-based detection
inline bool load_elenode(TetMeshData& m, const std::string& file) {
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

inline void save_elenode(const TetMeshData& m, const std::string& file) {
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
inline std::vector<std::array<int, 3>> boundary_faces(const std::vector<int>& tets) {
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

inline void save_obj(const std::vector<double>& verts, const std::vector<std::array<int, 3>>& faces, const std::string& file) {
    std::ofstream f(file);
    f << "# Generated by the MI355X WRAPD solver";
    for (size_t i = 0; i < verts.size() / 3; ++i)
        f << "\nv " << std::scientific << std::setprecision(15) << verts[3 * i] << ' ' << verts[3 * i + 1] << ' '
          << verts[3 * i + 2];
    for (const auto& t : faces) f << "\nf " << t[0] + 1 << ' ' << t[1] + 1 << ' ' << t[2] + 1;
    f << "\n";
}

// load_handles (samples/tests/quasi.cpp:146-159): whitespace-separated vertex indices
inline std::vector<int> load_handles(const std::string& file) {
    std::vector<int> pins;
    std::ifstream in(file);
    std::string line;
    while (std::getline(in, line)) {
        std::stringstream ss(line);
        int idx;
        while (ss >> idx) pins.push_back(idx);
    }
    return pins;
}

}  // namespace meshio
}  // namespace wrapd
