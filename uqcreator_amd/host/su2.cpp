// su2.cpp -- SU2 mesh ingestion for the flux sweeps: Mesh2D::LoadSU2MeshFromFile
// (mesh2d.cpp:68-255) with the cell geometry of triangle.cpp, quadrangle.cpp
// and cell.cpp, in the array layout of uqk_mesh.
//
// Kept as in the reference:
//   * nodes carry the id at the end of their line; elements are triangles
//     (type 5) or quadrilaterals (type 9) with their id last; cells keep the
//     file order and a neighbour is named by its element id;
//   * edge e of a cell joins its nodes e and e+1 (the last closes to node 0);
//     the normal of an edge is (-dy, dx) of A - B, flipped when it points to
//     the centroid, scaled by the edge length for the flux (triangle.cpp:31-70);
//   * a cell is a boundary cell when exactly two of its nodes lie on a marker
//     (triangle.cpp:4-13); its type is the type of the last marker element
//     (in node, marker, element order) whose first or second node is one of
//     the cell's (mesh2d.cpp:201-214); its boundary normal is the unit normal
//     of the edge with the largest neighbour id (cell.cpp:57-68);
//   * the centre is the vertex average, the triangle area the shoelace formula.
// Different on purpose:
//   * neighbours come from a hash of the node pair of every edge, O(edges),
//     instead of the all-pairs search of mesh2d.cpp:314-346 (the same result on
//     conforming meshes);
//   * maxEdge is the longest edge for quadrilaterals too (quadrangle.cpp:58,69
//     never update it), and the quadrilateral area is the shoelace formula
//     (equal to the reference's Bretschneider form up to rounding);
//   * element ids must be 0..NELEM-1 in file order (the reference indexes
//     cells by id), and a mesh mixes no triangles with quadrilaterals (every
//     cell of a uqk_mesh has the same number of edges).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "config.h"
#include "uqk.h"

namespace uqk_host {

namespace {

int boundary_type(const std::string& name) {
    if (name == "dirichlet") return UQK_BC_DIRICHLET;
    if (name == "noslip") return UQK_BC_NOSLIP;
    if (name == "swwall") return UQK_BC_SWWALL;
    // neumann and periodic have no code of their own in ExplicitEuler::Advance
    // (expliciteuler.cpp:20-72): their edges see uQ[nCells] like untyped ones
    if (name == "neumann" || name == "periodic") return UQK_BC_NONE;
    throw std::runtime_error("[mesh] invalid boundary condition '" + name + "'");
}

long trailing_number(const std::string& line) {
    const size_t p = line.find_first_of("0123456789");
    if (p == std::string::npos) throw std::runtime_error("[mesh] SU2: no count in '" + line + "'");
    return std::stol(line.substr(p));
}

struct Marker {
    std::string tag;
    int type = UQK_BC_NONE;
    std::vector<std::pair<long, long>> elems;  // first two node ids of each marker element
};

}  // namespace

Mesh Mesh::su2(const std::string& path, const std::vector<std::pair<std::string, std::string>>& bcs) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("[mesh] SU2 file not found: " + path);
    std::vector<std::string> lines;
    for (std::string l; std::getline(f, l);) lines.push_back(l);
    auto find = [&](const char* key) -> long {
        for (size_t i = 0; i < lines.size(); ++i)
            if (lines[i].find(key) != std::string::npos) return (long)i;
        return -1;
    };
    const long ldim = find("NDIME");
    if (ldim >= 0 && trailing_number(lines[ldim]) != 2) throw std::runtime_error("[mesh] SU2: NDIME must be 2");
    const long lpoin = find("NPOIN"), lelem = find("NELEM"), lmark = find("NMARK");
    if (lpoin < 0 || lelem < 0) throw std::runtime_error("[mesh] SU2: NPOIN or NELEM missing");

    // nodes: x y id
    const long np = trailing_number(lines[lpoin]);
    if (lpoin + np >= (long)lines.size()) throw std::runtime_error("[mesh] SU2: truncated NPOIN block");
    std::unordered_map<long, long> node_of;  // id -> index
    std::vector<long> node_id((size_t)np);
    std::vector<double> xy((size_t)np * 2);
    for (long i = 0; i < np; ++i) {
        std::istringstream ss(lines[lpoin + 1 + i]);
        double x, y;
        long id;
        if (!(ss >> x >> y >> id)) throw std::runtime_error("[mesh] SU2: bad node line " + std::to_string(i));
        xy[2 * i] = x;
        xy[2 * i + 1] = y;
        node_of[id] = i;
        node_id[i] = id;
    }
    auto node = [&](long id) -> long {
        auto it = node_of.find(id);
        if (it == node_of.end()) throw std::runtime_error("[mesh] SU2: unknown node id " + std::to_string(id));
        return it->second;
    };

    // markers: MARKER_TAG= t, MARKER_ELEMS= k, then k lines "3 a b"
    std::vector<Marker> markers;
    std::vector<char> on_boundary((size_t)np, 0);
    if (lmark >= 0) {
        const long nm = trailing_number(lines[lmark]);
        long l = lmark + 1;
        for (long m = 0; m < nm; ++m) {
            Marker mk;
            bool have_tag = false, have_elems = false;
            while ((!have_tag || !have_elems) && l < (long)lines.size()) {
                const std::string& s = lines[l++];
                if (s.find("MARKER_TAG") != std::string::npos) {
                    std::string t = s.substr(s.find('=') + 1);
                    t.erase(std::remove_if(t.begin(), t.end(), ::isspace), t.end());
                    mk.tag = t;
                    have_tag = true;
                } else if (s.find("MARKER_ELEMS") != std::string::npos) {
                    const long ne = trailing_number(s);
                    for (long e = 0; e < ne; ++e, ++l) {
                        if (l >= (long)lines.size()) throw std::runtime_error("[mesh] SU2: truncated marker " + mk.tag);
                        std::istringstream ss(lines[l]);
                        long type, a, b;
                        if (!(ss >> type >> a >> b)) throw std::runtime_error("[mesh] SU2: bad marker element");
                        on_boundary[node(a)] = 1;
                        on_boundary[node(b)] = 1;
                        mk.elems.push_back({a, b});
                    }
                    have_elems = true;
                } else if (!s.empty() && s.find_first_not_of(" \t\r") != std::string::npos) {
                    throw std::runtime_error("[mesh] SU2: unexpected line in marker block: " + s);
                }
            }
            bool typed = false;
            for (const auto& bc : bcs)
                if (bc.first == mk.tag) {
                    mk.type = boundary_type(bc.second);
                    typed = true;
                }
            if (!typed) throw std::runtime_error("[mesh] SU2: marker '" + mk.tag + "' has no entry in SU2BC");
            markers.push_back(mk);
        }
    }

    // elements
    const long nc = trailing_number(lines[lelem]);
    if (lelem + nc >= (long)lines.size()) throw std::runtime_error("[mesh] SU2: truncated NELEM block");
    Mesh m;
    m.n_cells = nc;
    m.dim = 2;
    std::vector<long> cn;  // cell nodes (indices), E per cell
    int E = 0;
    for (long c = 0; c < nc; ++c) {
        std::istringstream ss(lines[lelem + 1 + c]);
        int type;
        if (!(ss >> type)) throw std::runtime_error("[mesh] SU2: bad element line " + std::to_string(c));
        const int k = type == 5 ? 3 : type == 9 ? 4 : 0;
        if (!k) throw std::runtime_error("[mesh] SU2: unsupported element type " + std::to_string(type));
        if (E && k != E) throw std::runtime_error("[mesh] SU2: triangles and quadrilaterals mixed");
        E = k;
        long ids[4], id;
        for (int i = 0; i < k; ++i)
            if (!(ss >> ids[i])) throw std::runtime_error("[mesh] SU2: bad element line " + std::to_string(c));
        if (!(ss >> id) || id != c) throw std::runtime_error("[mesh] SU2: element ids must be 0..NELEM-1 in order");
        for (int i = 0; i < k; ++i) cn.push_back(node(ids[i]));
    }
    m.n_edges = E;
    m.cell_node_count = E;
    m.cell_nodes.assign(cn.begin(), cn.end());
    m.node_xy = xy;
    const size_t N = (size_t)nc;
    m.nbr.assign(N * E, (int32_t)nc);
    m.unit_normal.assign(N * E * 2, 0.0);
    m.normal.assign(N * E * 2, 0.0);
    m.area.assign(N, 0.0);
    m.dx_param.assign(N, 0.0);
    m.bc.assign(N, UQK_BC_NONE);
    m.bnormal.assign(N * 2, 0.0);
    m.center.assign(N * 2, 0.0);

    // neighbours: edge (node pair) -> first cell and local edge seen
    std::unordered_map<uint64_t, std::pair<long, int>> edge_owner;
    edge_owner.reserve(N * E);
    for (long c = 0; c < nc; ++c) {
        const long* v = &cn[(size_t)c * E];
        for (int e = 0; e < E; ++e) {
            long a = v[e], b = v[(e + 1) % E];
            if (a > b) std::swap(a, b);
            const uint64_t key = ((uint64_t)a << 32) | (uint64_t)b;
            auto it = edge_owner.find(key);
            if (it == edge_owner.end()) {
                edge_owner.emplace(key, std::make_pair(c, e));
            } else {
                const long o = it->second.first;
                const int oe = it->second.second;
                m.nbr[(size_t)o * E + oe] = (int32_t)c;
                m.nbr[(size_t)c * E + e] = (int32_t)o;
            }
        }
    }

    for (long c = 0; c < nc; ++c) {
        const long* v = &cn[(size_t)c * E];
        double cx = 0.0, cy = 0.0;
        for (int i = 0; i < E; ++i) {
            cx += xy[2 * v[i]];
            cy += xy[2 * v[i] + 1];
        }
        cx /= E;
        cy /= E;
        m.center[2 * c] = cx;
        m.center[2 * c + 1] = cy;
        double twice = 0.0, max_edge = 0.0;
        for (int e = 0; e < E; ++e) {
            const long A = v[e], B = v[(e + 1) % E];
            const double ax = xy[2 * A], ay = xy[2 * A + 1], bx = xy[2 * B], by = xy[2 * B + 1];
            twice += ax * by - bx * ay;
            const double dx = ax - bx, dy = ay - by;
            const double len = std::sqrt(dx * dx + dy * dy);
            max_edge = std::max(max_edge, len);
            double nx = -dy, ny = dx;
            if (nx * (cx - ax) + ny * (cy - ay) > 0.0) {
                nx = -nx;
                ny = -ny;
            }
            const double nn = std::sqrt(nx * nx + ny * ny);
            nx /= nn;
            ny /= nn;
            const size_t o = ((size_t)c * E + e) * 2;
            m.unit_normal[o] = nx;
            m.unit_normal[o + 1] = ny;
            m.normal[o] = len * nx;
            m.normal[o + 1] = len * ny;
        }
        m.area[c] = std::fabs(twice) / 2.0;
        m.dx_param[c] = max_edge / m.area[c];
        int nb = 0;
        for (int i = 0; i < E; ++i) nb += on_boundary[v[i]];
        if (nb == 2) {
            int type = UQK_BC_NONE;
            for (int i = 0; i < E; ++i) {
                const long vid = node_id[v[i]];
                for (const Marker& mk : markers)
                    for (const auto& el : mk.elems)
                        if (el.first == vid || el.second == vid) type = mk.type;
            }
            m.bc[c] = type;
            bool open_edge = false;
            for (int e = 0; e < E; ++e) open_edge = open_edge || m.nbr[(size_t)c * E + e] == (int32_t)nc;
            if (!open_edge)  // mesh2d.cpp:222-253 stops here too
                throw std::runtime_error("[mesh] SU2: wrong boundary cell " + std::to_string(c) +
                                         " (two marker nodes, no boundary edge)");
            int id = 0;
            for (int e = 1; e < E; ++e)
                if (m.nbr[(size_t)c * E + e] > m.nbr[(size_t)c * E + id]) id = e;
            m.bnormal[2 * c] = m.unit_normal[((size_t)c * E + id) * 2];
            m.bnormal[2 * c + 1] = m.unit_normal[((size_t)c * E + id) * 2 + 1];
        }
    }
    return m;
}

}  // namespace uqk_host

extern "C" int64_t uqk_su2_mesh(const char* path, const char* bc_spec, int32_t* n_edges, int32_t* nbr,
                                double* unit_normal, double* normal, double* area, double* dx_param, int32_t* bc,
                                double* bnormal, double* center, char* err, int32_t errlen) {
    try {
        if (!path) throw std::runtime_error("null path");
        std::vector<std::pair<std::string, std::string>> bcs;
        std::string spec = bc_spec ? bc_spec : "";
        std::stringstream ss(spec);
        for (std::string item; std::getline(ss, item, ';');) {
            if (item.empty()) continue;
            const size_t eq = item.find('=');
            if (eq == std::string::npos) throw std::runtime_error("bc_spec items are tag=type");
            bcs.push_back({item.substr(0, eq), item.substr(eq + 1)});
        }
        const uqk_host::Mesh m = uqk_host::Mesh::su2(path, bcs);
        if (n_edges) *n_edges = m.n_edges;
        auto put = [](auto* dst, const auto& src) {
            if (dst) std::copy(src.begin(), src.end(), dst);
        };
        put(nbr, m.nbr);
        put(unit_normal, m.unit_normal);
        put(normal, m.normal);
        put(area, m.area);
        put(dx_param, m.dx_param);
        put(bc, m.bc);
        put(bnormal, m.bnormal);
        put(center, m.center);
        return (int64_t)m.n_cells;
    } catch (const std::exception& e) {
        if (err && errlen > 0) {
            std::strncpy(err, e.what(), (size_t)errlen - 1);
            err[errlen - 1] = '\0';
        }
        return -1;
    }
}
