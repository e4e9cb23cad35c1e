// Mesh inputs of the C++ API (SdfSurfaceMesh.h:16-24 of the reference):
// the OBJ reader and the NEXDYN_CONTACT_AWARE_SURFACE 1 reader, with the
// reference's parsing rules and error texts (SdfSurfaceMesh.cpp:111-344).
// Host text parsing; the result feeds SdfBuilder.
#pragma once

#include <array>
#include <cmath>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "SdfOracle/Types.h"

namespace sdf {

struct TriangleMesh
{
    std::vector<Vector3d> vertices;
    std::vector<std::array<int, 3>> triangles;
};

struct CornerNormalTriangleMesh
{
    std::vector<Vector3d> vertices;
    std::vector<std::array<int, 3>> triangles;
    std::vector<std::array<Vector3d, 3>> cornerNormals;
};

// "v x y z" lines and the first three indices of "f" lines (v, v/vt, v/vt/vn,
// v//vn); further face indices are ignored.
inline TriangleMesh loadObjFile(const std::string& path)
{
    std::ifstream in(path);
    if (!in.is_open()) throw std::runtime_error("Cannot open OBJ file: " + path);
    TriangleMesh mesh;
    std::string line;
    while (std::getline(in, line))
    {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream ss(line);
        std::string token;
        ss >> token;
        if (token == "v")
        {
            double x = 0.0, y = 0.0, z = 0.0;
            ss >> x >> y >> z;
            mesh.vertices.emplace_back(x, y, z);
        }
        else if (token == "f")
        {
            std::array<int, 3> tri{};
            for (int c = 0; c < 3; ++c)
            {
                std::string t;
                ss >> t;
                tri[size_t(c)] = std::stoi(t.substr(0, t.find('/'))) - 1;
            }
            mesh.triangles.push_back(tri);
        }
    }
    return mesh;
}

namespace detail {

struct SurfaceState
{
    bool active = false;
    std::string name;
    int expectedVertices = -1, expectedTriangles = -1;
    int seenVertices = 0, seenTriangles = 0;
    double T[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
    double N[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}; // inverse transpose of T's linear part
    std::unordered_map<int, int> ids;
};

inline std::runtime_error surfaceError(const std::string& path, int ln, const std::string& msg)
{
    return std::runtime_error(path + ":" + std::to_string(ln) + ": " + msg);
}

inline std::string stripComment(std::string line)
{
    const size_t c = line.find('#');
    if (c != std::string::npos) line.erase(c);
    const size_t a = line.find_first_not_of(" \t\r\n");
    if (a == std::string::npos) return {};
    const size_t b = line.find_last_not_of(" \t\r\n");
    return line.substr(a, b - a + 1);
}

inline void closeSurface(const std::string& path, int ln, SurfaceState& s)
{
    if (!s.active) return;
    if (s.expectedVertices < 0) throw surfaceError(path, ln, "surface '" + s.name + "' is missing vertices block");
    if (s.expectedTriangles < 0) throw surfaceError(path, ln, "surface '" + s.name + "' is missing triangles block");
    if (s.seenVertices != s.expectedVertices) throw surfaceError(path, ln, "surface '" + s.name + "' vertex count mismatch");
    if (s.seenTriangles != s.expectedTriangles)
        throw surfaceError(path, ln, "surface '" + s.name + "' triangle count mismatch");
    s = SurfaceState{};
}

} // namespace detail

inline CornerNormalTriangleMesh loadContactAwareSurfaceFile(const std::string& path)
{
    using detail::surfaceError;
    std::ifstream in(path);
    if (!in.is_open()) throw std::runtime_error("Cannot open contact-aware surface file: " + path);
    CornerNormalTriangleMesh mesh;
    detail::SurfaceState st;
    bool header = false;
    std::string raw;
    int ln = 0;
    while (std::getline(in, raw))
    {
        ++ln;
        const std::string line = detail::stripComment(raw);
        if (line.empty()) continue;
        std::istringstream ss(line);
        std::string token;
        ss >> token;
        if (!header)
        {
            int version = 0;
            if (token != "NEXDYN_CONTACT_AWARE_SURFACE" || !(ss >> version) || version != 1)
                throw surfaceError(path, ln, "expected header 'NEXDYN_CONTACT_AWARE_SURFACE 1'");
            header = true;
            continue;
        }
        if (token == "surface")
        {
            detail::closeSurface(path, ln, st);
            st.active = true;
            if (!(ss >> st.name) || st.name.empty()) throw surfaceError(path, ln, "surface requires a name");
            continue;
        }
        if (token == "end")
        {
            detail::closeSurface(path, ln, st);
            continue;
        }
        if (!st.active) throw surfaceError(path, ln, "expected surface block");
        if (token == "transform")
        {
            for (int r = 0; r < 4; ++r)
                for (int c = 0; c < 4; ++c)
                    if (!(ss >> st.T[r][c])) throw surfaceError(path, ln, "transform requires 16 row-major values");
            const double(&L)[4][4] = st.T;
            const double det = L[0][0] * (L[1][1] * L[2][2] - L[1][2] * L[2][1]) -
                               L[0][1] * (L[1][0] * L[2][2] - L[1][2] * L[2][0]) +
                               L[0][2] * (L[1][0] * L[2][1] - L[1][1] * L[2][0]);
            if (std::abs(det) <= 1.0e-15) throw surfaceError(path, ln, "transform linear part must be invertible");
            // (L^-1)^T = cofactor(L) / det
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c)
                {
                    const int r1 = (r + 1) % 3, r2 = (r + 2) % 3, c1 = (c + 1) % 3, c2 = (c + 2) % 3;
                    st.N[r][c] = (L[r1][c1] * L[r2][c2] - L[r1][c2] * L[r2][c1]) / det;
                }
            continue;
        }
        if (token == "vertices")
        {
            if (!(ss >> st.expectedVertices) || st.expectedVertices < 0)
                throw surfaceError(path, ln, "vertices requires a non-negative count");
            continue;
        }
        if (token == "triangles")
        {
            if (!(ss >> st.expectedTriangles) || st.expectedTriangles < 0)
                throw surfaceError(path, ln, "triangles requires a non-negative count");
            continue;
        }
        if (token == "v")
        {
            if (st.expectedVertices < 0) throw surfaceError(path, ln, "vertex appears before vertices block");
            int id = 0;
            double x, y, z;
            if (!(ss >> id >> x >> y >> z)) throw surfaceError(path, ln, "vertex requires: v <id> <x> <y> <z>");
            if (st.ids.count(id)) throw surfaceError(path, ln, "duplicate vertex id in surface");
            st.ids.emplace(id, int(mesh.vertices.size()));
            double h[4];
            for (int r = 0; r < 4; ++r) h[r] = st.T[r][0] * x + st.T[r][1] * y + st.T[r][2] * z + st.T[r][3];
            if (std::abs(h[3]) > 1.0e-15)
                mesh.vertices.emplace_back(h[0] / h[3], h[1] / h[3], h[2] / h[3]);
            else
                mesh.vertices.emplace_back(h[0], h[1], h[2]);
            ++st.seenVertices;
            continue;
        }
        if (token == "tri")
        {
            if (st.expectedTriangles < 0) throw surfaceError(path, ln, "triangle appears before triangles block");
            int tid = 0, vid[3];
            double nv[3][3];
            if (!(ss >> tid >> vid[0] >> vid[1] >> vid[2] >> nv[0][0] >> nv[0][1] >> nv[0][2] >> nv[1][0] >>
                  nv[1][1] >> nv[1][2] >> nv[2][0] >> nv[2][1] >> nv[2][2]))
                throw surfaceError(path, ln, "triangle requires id, 3 vertex ids, and 3 corner normals");
            std::array<int, 3> tri{};
            std::array<Vector3d, 3> cn;
            for (int c = 0; c < 3; ++c)
            {
                const auto it = st.ids.find(vid[c]);
                if (it == st.ids.end()) throw surfaceError(path, ln, "triangle references unknown vertex id");
                tri[size_t(c)] = it->second;
                double n[3];
                for (int r = 0; r < 3; ++r) n[r] = st.N[r][0] * nv[c][0] + st.N[r][1] * nv[c][1] + st.N[r][2] * nv[c][2];
                const double len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
                if (!std::isfinite(n[0]) || !std::isfinite(n[1]) || !std::isfinite(n[2]) || !std::isfinite(len) ||
                    len <= 1.0e-15)
                    throw surfaceError(path, ln, "corner normal must be a finite non-zero vector");
                cn[size_t(c)] = Vector3d(n[0] / len, n[1] / len, n[2] / len);
            }
            const Vector3d& a = mesh.vertices[size_t(tri[0])];
            const Vector3d e0 = mesh.vertices[size_t(tri[1])] - a, e1 = mesh.vertices[size_t(tri[2])] - a;
            if (e0.cross(e1).norm() <= 1.0e-15) throw surfaceError(path, ln, "triangle is degenerate");
            mesh.triangles.push_back(tri);
            mesh.cornerNormals.push_back(cn);
            ++st.seenTriangles;
            continue;
        }
        throw surfaceError(path, ln, "unknown token '" + token + "'");
    }
    if (!header) throw std::runtime_error(path + ": missing contact-aware surface header");
    detail::closeSurface(path, ln, st);
    if (mesh.vertices.empty() || mesh.triangles.empty())
        throw std::runtime_error(path + ": contact-aware surface contains no triangles");
    return mesh;
}

} // namespace sdf
