// Mesh readers and the structured Kuhn-beam generator of the drop-in API.
#include "../../../include/tlfea/utils/mesh_io.h"

#include <algorithm>
#include <array>
#include <climits>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <tuple>
#include <vector>

namespace tlfea {

namespace {
void err(const std::string& m) { std::fprintf(stderr, "[TLFEA ERROR] %s\n", m.c_str()); }
}  // namespace

// TetGen order [v0 v1 v2 v3 (3-4) (1-4) (1-2) (2-3) (2-4) (1-3)] -> standard
// [v0 v1 v2 v3 (0-1) (1-2) (0-2) (0-3) (1-3) (2-3)] (cpu_utils.cc:596-611).
void FEAT10_remap_tetgen_indices(const VectorXi& tetgen_elem, VectorXi& standard_elem) {
    if (tetgen_elem.size() != 10 || standard_elem.size() != 10) {
        err("Element arrays must have size 10 for T10 elements");
        return;
    }
    static const int from[10] = {0, 1, 2, 3, 6, 7, 9, 5, 8, 4};
    for (int i = 0; i < 10; i++) standard_elem(i) = tetgen_elem(from[i]);
}

int FEAT10_read_nodes(const std::string& filename, MatrixXR& nodes) {
    std::ifstream f(filename);
    if (!f) {
        err("Could not open node file " + filename);
        return 0;
    }
    std::string line;
    if (!std::getline(f, line)) {
        err("Could not read header line from " + filename);
        return 0;
    }
    int n = 0, dim = 0;
    std::istringstream(line) >> n >> dim;
    if (dim != 3) {
        err("Only 3D nodes are supported");
        return 0;
    }
    nodes.resize(n, 3);
    nodes.setZero();
    std::vector<std::tuple<int, Real, Real, Real>> rows;
    int min_id = INT_MAX;
    for (int i = 0; i < n && std::getline(f, line);) {
        if (line.empty()) continue;
        std::istringstream s(line);
        int id;
        Real x, y, z;
        if (s >> id >> x >> y >> z) {
            rows.emplace_back(id, x, y, z);
            min_id = std::min(min_id, id);
        }
        i++;
    }
    const int off = (min_id == 0) ? 0 : 1;
    for (auto& r : rows) {
        const int idx = std::get<0>(r) - off;
        if (idx >= 0 && idx < n) {
            nodes(idx, 0) = std::get<1>(r);
            nodes(idx, 1) = std::get<2>(r);
            nodes(idx, 2) = std::get<3>(r);
        }
    }
    return n;
}

int FEAT10_read_elements(const std::string& filename, MatrixXi& elements) {
    std::ifstream f(filename);
    if (!f) {
        err("Could not open element file " + filename);
        return 0;
    }
    std::string line;
    if (!std::getline(f, line)) {
        err("Could not read header line from " + filename);
        return 0;
    }
    int n = 0, npe = 0;
    std::istringstream(line) >> n >> npe;
    if (npe != 10) {
        err("Only T10 elements (10 nodes) are supported");
        return 0;
    }
    elements.resize(n, 10);
    elements.setZero();
    std::vector<std::pair<int, std::array<int, 10>>> recs;
    int min_e = INT_MAX, min_n = INT_MAX;
    for (int i = 0; i < n && std::getline(f, line);) {
        if (line.empty()) continue;
        std::istringstream s(line);
        int id;
        std::array<int, 10> v{};
        s >> id;
        for (int j = 0; j < 10; j++)
            if (s >> v[j]) min_n = std::min(min_n, v[j]);
        min_e = std::min(min_e, id);
        recs.emplace_back(id, v);
        i++;
    }
    const int eoff = (min_e == 0) ? 0 : 1, noff = (min_n == 0) ? 0 : 1;
    VectorXi tg(10), st(10);
    for (auto& r : recs) {
        for (int j = 0; j < 10; j++) tg(j) = r.second[j] - noff;
        FEAT10_remap_tetgen_indices(tg, st);
        const int idx = r.first - eoff;
        if (idx >= 0 && idx < n)
            for (int j = 0; j < 10; j++) elements(idx, j) = st(j);
    }
    return n;
}

namespace {
// text of the first <DataArray ...> whose attributes contain `key`, within [from, to)
bool data_array(const std::string& doc, const std::string& key, size_t from, size_t to, std::string& out) {
    size_t p = from;
    while ((p = doc.find("<DataArray", p)) != std::string::npos && p < to) {
        const size_t gt = doc.find('>', p);
        if (gt == std::string::npos) return false;
        if (doc.substr(p, gt - p).find(key) != std::string::npos) {
            const size_t end = doc.find("</DataArray>", gt);
            std::string body = doc.substr(gt + 1, end - gt - 1);
            // drop nested tags such as <InformationKey ...>
            std::string clean;
            bool in_tag = false;
            for (char c : body) {
                if (c == '<') in_tag = true;
                if (!in_tag) clean.push_back(c);
                if (c == '>') in_tag = false;
            }
            out = clean;
            return true;
        }
        p = gt;
    }
    return false;
}
}  // namespace

int LoadVtuTet(const std::string& filename, MatrixXR& nodes, MatrixXi& elements) {
    std::ifstream f(filename);
    if (!f) {
        err("Could not open VTU file " + filename);
        return 0;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string doc = ss.str();
    auto attr = [&](const std::string& name) -> long {
        const size_t p = doc.find(name + "=\"");
        if (p == std::string::npos) return -1;
        return std::atol(doc.c_str() + p + name.size() + 2);
    };
    const long npts = attr("NumberOfPoints"), ncell = attr("NumberOfCells");
    const size_t pts = doc.find("<Points>"), cells = doc.find("<Cells>");
    std::string txt;
    if (npts <= 0 || ncell <= 0 || pts == std::string::npos || cells == std::string::npos ||
        !data_array(doc, "", pts, cells, txt)) {
        err("unsupported VTU (ASCII UnstructuredGrid with Points and Cells expected)");
        return 0;
    }
    nodes.resize(npts, 3);
    {
        std::istringstream s(txt);
        for (long i = 0; i < npts; i++) s >> nodes(i, 0) >> nodes(i, 1) >> nodes(i, 2);
    }
    std::string conn, types;
    if (!data_array(doc, "Name=\"connectivity\"", cells, doc.size(), conn) ||
        !data_array(doc, "Name=\"types\"", cells, doc.size(), types)) {
        err("VTU file without connectivity/types arrays");
        return 0;
    }
    std::istringstream ts(types);
    int t0 = 0, t;
    ts >> t0;
    while (ts >> t)
        if (t != t0) {
            err("mixed cell types are not supported");
            return 0;
        }
    const int nn = (t0 == 24) ? 10 : (t0 == 10 ? 4 : 0);
    if (nn == 0) {
        err("only VTK_QUADRATIC_TETRA (24) and VTK_TETRA (10) cells are supported");
        return 0;
    }
    elements.resize(ncell, nn);
    std::istringstream cs(conn);
    for (long e = 0; e < ncell; e++)
        for (int a = 0; a < nn; a++) {
            long v;
            cs >> v;
            elements(e, a) = (int)v;
        }
    return (int)ncell;
}

void GenerateKuhnBeam(int nx, int ny, int nz, Real lx, Real ly, Real lz, bool tet10, MatrixXR& nodes,
                      MatrixXi& elements) {
    const int s = tet10 ? 2 : 1;
    const int NX = s * nx + 1, NY = s * ny + 1, NZ = s * nz + 1;
    nodes.resize((std::ptrdiff_t)NX * NY * NZ, 3);
    for (int I = 0; I < NX; I++)
        for (int J = 0; J < NY; J++)
            for (int K = 0; K < NZ; K++) {
                const std::ptrdiff_t id = ((std::ptrdiff_t)I * NY + J) * NZ + K;
                nodes(id, 0) = I * (lx / (s * nx));
                nodes(id, 1) = J * (ly / (s * ny));
                nodes(id, 2) = K * (lz / (s * nz));
            }
    // six Kuhn paths 0 -> e_p0 -> e_p0+e_p1 -> (1,1,1); odd permutations get corners 1 and 2
    // swapped so every tet has positive signed volume
    static const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    static const int edges[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}};
    const int nn = tet10 ? 10 : 4;
    elements.resize((std::ptrdiff_t)nx * ny * nz * 6, nn);
    std::ptrdiff_t e = 0;
    for (int ci = 0; ci < nx; ci++)
        for (int cj = 0; cj < ny; cj++)
            for (int ck = 0; ck < nz; ck++)
                for (auto& p : perms) {
                    int c[4][3] = {{0, 0, 0}};
                    for (int k = 1; k < 4; k++) {
                        for (int d = 0; d < 3; d++) c[k][d] = c[k - 1][d];
                        c[k][p[k - 1]] += 1;
                    }
                    const bool odd = !((p[0] == 0 && p[1] == 1) || (p[0] == 1 && p[1] == 2) || (p[0] == 2 && p[1] == 0));
                    if (odd)
                        for (int d = 0; d < 3; d++) std::swap(c[1][d], c[2][d]);
                    int lat[4][3];
                    for (int k = 0; k < 4; k++) {
                        lat[k][0] = s * (ci + c[k][0]);
                        lat[k][1] = s * (cj + c[k][1]);
                        lat[k][2] = s * (ck + c[k][2]);
                    }
                    auto id = [&](int a, int b, int cc) { return (a * NY + b) * NZ + cc; };
                    for (int k = 0; k < 4; k++) elements(e, k) = id(lat[k][0], lat[k][1], lat[k][2]);
                    if (tet10)
                        for (int k = 0; k < 6; k++) {
                            const int a = edges[k][0], b = edges[k][1];
                            elements(e, 4 + k) = id((lat[a][0] + lat[b][0]) / 2, (lat[a][1] + lat[b][1]) / 2,
                                                    (lat[a][2] + lat[b][2]) / 2);
                        }
                    e++;
                }
}

// Legacy ASCII VTK of a tetrahedral mesh at deformed positions, with the displacement as
// point data: the layout of WriteFEAT10ToVTK / WriteFEAT4ToVTK (mesh_utils.cc:962-1077).
// Values go through the stream's default formatting (6 significant digits), as there.
static bool write_tet_vtk(const std::string& filename, int nn, const char* title, const MatrixXR& nodes,
                          const MatrixXi& elements, const VectorXR& x, const VectorXR& y, const VectorXR& z) {
    std::ofstream file(filename);
    if (!file.is_open()) {
        err("Failed to open file for writing: " + filename);
        return false;
    }
    const std::ptrdiff_t n_nodes = nodes.rows(), n_el = elements.rows();
    if (elements.cols() != nn || nodes.cols() < 3 || x.size() != n_nodes || y.size() != n_nodes ||
        z.size() != n_nodes) {
        err("VTK writer: array sizes do not match the mesh");
        return false;
    }
    file << "# vtk DataFile Version 3.0\n" << title << "\nASCII\nDATASET UNSTRUCTURED_GRID\n";
    file << "POINTS " << n_nodes << " double\n";
    for (std::ptrdiff_t i = 0; i < n_nodes; i++) file << x(i) << " " << y(i) << " " << z(i) << "\n";
    file << "\nCELLS " << n_el << " " << n_el * (nn + 1) << "\n";
    for (std::ptrdiff_t e = 0; e < n_el; e++) {
        file << nn;
        for (int a = 0; a < nn; a++) file << " " << elements(e, a);
        file << "\n";
    }
    file << "\nCELL_TYPES " << n_el << "\n";
    const char* ctype = nn == 10 ? "24\n" : "10\n";  // VTK_QUADRATIC_TETRA / VTK_TETRA
    for (std::ptrdiff_t e = 0; e < n_el; e++) file << ctype;
    file << "\nPOINT_DATA " << n_nodes << "\nVECTORS displacement double\n";
    for (std::ptrdiff_t i = 0; i < n_nodes; i++)
        file << x(i) - nodes(i, 0) << " " << y(i) - nodes(i, 1) << " " << z(i) - nodes(i, 2) << "\n";
    return true;
}

namespace ANCFCPUUtils {
bool WriteFEAT10ToVTK(const std::string& filename, const MatrixXR& nodes, const MatrixXi& elements, const VectorXR& x,
                      const VectorXR& y, const VectorXR& z) {
    return write_tet_vtk(filename, 10, "FEAT10 Tetrahedral Mesh", nodes, elements, x, y, z);
}
bool WriteFEAT4ToVTK(const std::string& filename, const MatrixXR& nodes, const MatrixXi& elements, const VectorXR& x,
                     const VectorXR& y, const VectorXR& z) {
    return write_tet_vtk(filename, 4, "FEAT4 Tetrahedral Mesh", nodes, elements, x, y, z);
}
}  // namespace ANCFCPUUtils

}  // namespace tlfea
