// Command-line driver over the drop-in C++ API (FEASolver, GPU_FEAT10_Data / GPU_FEAT4_Data,
// LinearStaticSolver, SyncedNesterovSolver, SyncedAdamW[Nocoop]Solver).  It runs the
// reference's example scenarios (examples/beam_linear_static.cc, beam_linear_static_t4.cc,
// test_feat10_nesterov.cc, test_feat10_adamw.cc) and writes the final nodal positions as
// raw float64 (x block, y block, z block) to --out, plus a JSON summary on stdout.
//
//   tlfea_driver linear_static  --mesh kuhn:20,2,2 | kuhn4:12,3,2 | vtu:PATH | tetgen:PREFIX
//                 [--E 2.1e11 --nu 0.3 --load -1000 --tol 1e-10 --max-iter 50000]
//                 [--assembled 1: assemble K as the reference does] [--solves N]
//                 [--two-level DEG: FP32 two-level preconditioned CG] --out FILE
//   tlfea_driver nesterov|adamw|adamw_nocoop --mesh tetgen:PREFIX|kuhn:.. [--steps 3] [--E ..]
//                 [--rho0 ..] [--fix z0|x0] [--node-load NODE,FX] --out FILE
//                 [--update-fixed-after K --fix2 SPEC]: before step K, UpdateNodalFixed to the
//                 SPEC set (x0 | y0 | z0 | x0+z0, ":rev" reverses the order), CalcConstraintData,
//                 ConvertToCSR_ConstraintJacT (FEAT10Data.cu:674-751)
//                 [--jac-out FILE]: the J and J^T CSR (ConvertToCSR_ConstraintJac/JacT) as text
//                 [--load-file F]: f_ext from raw float64 [3N]; [--start-file F]: positions
//                 (raw float64 x, y, z blocks) set by UpdatePositions after CalcDnDuPre, so the
//                 constraint targets and rest geometry stay the mesh's
//   any mode: [--mr MU10,MU01,KAPPA] switches the element to Mooney-Rivlin (SetMooneyRivlin)
//             [--vtk FILE] WriteOutputVTK; [--vtk-feat FILE] ANCFCPUUtils::WriteFEAT10ToVTK /
//             WriteFEAT4ToVTK of the final positions against the input nodes
//   tlfea_driver write_vtk --mesh SPEC --vtk-feat FILE [--shear S]: no device work; the mesh
//             at x + S (y z, -x z, x y) through WriteFEAT10ToVTK / WriteFEAT4ToVTK
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <string>
#include <type_traits>
#include <vector>

#include "../../../include/tlfea/FEASolver.h"
#include "../../../include/tlfea/solvers/LinearStaticSolver.h"
#include "../../../include/tlfea/solvers/SyncedAdamWNocoop.h"
#include "../../../include/tlfea/utils/mesh_io.h"
#include "../../../include/tlfea/utils/quadrature_utils.h"

using namespace tlfea;

namespace {
std::map<std::string, std::string> parse(int argc, char** argv) {
    std::map<std::string, std::string> a;
    for (int i = 2; i + 1 < argc; i += 2) a[argv[i]] = argv[i + 1];
    return a;
}
// a double in JSON without losing small magnitudes (std::to_string prints %f)
std::string sci(double v) {
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%.6e", v);
    return buf;
}
double num(std::map<std::string, std::string>& a, const char* k, double d) {
    return a.count(k) ? std::atof(a[k].c_str()) : d;
}

bool load_mesh(const std::string& spec, MatrixXR& nodes, MatrixXi& elems) {
    if (spec.rfind("kuhn", 0) == 0) {
        const bool t4 = spec.rfind("kuhn4:", 0) == 0;
        int nx, ny, nz;
        if (std::sscanf(spec.c_str() + (t4 ? 6 : 5), "%d,%d,%d", &nx, &ny, &nz) != 3) return false;
        GenerateKuhnBeam(nx, ny, nz, 1.0, 0.1, 0.1, !t4, nodes, elems);
        return true;
    }
    if (spec.rfind("vtu:", 0) == 0) return LoadVtuTet(spec.substr(4), nodes, elems) > 0;
    if (spec.rfind("tetgen:", 0) == 0) {
        const std::string p = spec.substr(7);
        return FEAT10_read_nodes(p + ".node", nodes) > 0 && FEAT10_read_elements(p + ".ele", elems) > 0;
    }
    return false;
}

// Node selections of the scenarios: nodes on the x = x_min ("x0"), y = 0 or z = 0 face, or
// on any of several faces ("x0+z0"), in ascending order; a ":rev" suffix reverses it.
std::vector<int> select_nodes(const std::string& spec, const VectorXR& hx, const VectorXR& hy, const VectorXR& hz) {
    const bool rev = spec.size() > 4 && spec.compare(spec.size() - 4, 4, ":rev") == 0;
    const std::string faces = "+" + (rev ? spec.substr(0, spec.size() - 4) : spec) + "+";
    auto has = [&](const char* f) { return faces.find(std::string("+") + f + "+") != std::string::npos; };
    const Real xmin = hx.minCoeff(), L = hx.maxCoeff() - xmin;
    std::vector<int> out;
    for (std::ptrdiff_t i = 0; i < hx.size(); i++) {
        if ((has("z0") && std::abs(hz(i)) < 1e-8) || (has("y0") && std::abs(hy(i)) < 1e-8) ||
            (has("x0") && std::abs(hx(i) - xmin) < 5e-4 * L))
            out.push_back((int)i);
    }
    if (rev) std::reverse(out.begin(), out.end());
    return out;
}

VectorXi to_vectorxi(const std::vector<int>& v) {
    VectorXi r((std::ptrdiff_t)v.size());
    for (size_t k = 0; k < v.size(); k++) r((std::ptrdiff_t)k) = v[k];
    return r;
}

template <typename T>
void write_array(std::ofstream& f, const char* name, const std::vector<T>& v) {
    f << name << " " << v.size();
    char buf[40];
    for (const T& x : v) {
        std::snprintf(buf, sizeof(buf), std::is_integral<T>::value ? " %.0f" : " %.17g", (double)x);
        f << buf;
    }
    f << "\n";
}

template <typename TData>
void write_jacobians(const std::string& path, TData* b) {
    std::ofstream f(path);
    for (int t = 0; t < 2; t++) {
        std::vector<int> off, col;
        std::vector<Real> val;
        b->RetrieveConstraintJacobianCSRToCPU(t == 1, off, col, val);
        const std::string p = t ? "jt_" : "j_";
        write_array(f, (p + "offsets").c_str(), off);
        write_array(f, (p + "columns").c_str(), col);
        write_array(f, (p + "values").c_str(), val);
    }
}

void write_positions(const std::string& path, const VectorXR& x, const VectorXR& y, const VectorXR& z) {
    std::ofstream f(path, std::ios::binary);
    for (const VectorXR* v : {&x, &y, &z}) f.write(reinterpret_cast<const char*>(v->data()), sizeof(double) * v->size());
}

template <typename TData>
int run(const std::string& mode, std::map<std::string, std::string>& a, MatrixXR& nodes, MatrixXi& elems) {
    const int n_nodes = (int)nodes.rows(), n_el = (int)elems.rows();
    VectorXR hx(n_nodes), hy(n_nodes), hz(n_nodes);
    for (int i = 0; i < n_nodes; i++) {
        hx(i) = nodes(i, 0);
        hy(i) = nodes(i, 1);
        hz(i) = nodes(i, 2);
    }
    TData body(n_el, n_nodes);
    FEASolver fea;
    fea.AddElement("body", &body);
    auto* b = static_cast<TData*>(fea.GetElement("body"));
    b->Initialize();
    const bool t10 = TData::N_NODES_PER_ELEM == 10;
    // boundary conditions and loads, as the reference drivers set them
    VectorXR f(3 * n_nodes);
    f.setZero();
    const std::string fix = a.count("--fix") ? a["--fix"] : (mode == "linear_static" ? "x0" : "z0");
    const Real xmax = hx.maxCoeff(), L = xmax - hx.minCoeff();
    const std::vector<int> fixed = select_nodes(fix, hx, hy, hz);
    if (a.count("--load-file")) {
        std::ifstream lf(a["--load-file"], std::ios::binary);
        lf.read(reinterpret_cast<char*>(f.data()), sizeof(double) * 3 * (size_t)n_nodes);
        if (!lf) {
            std::fprintf(stderr, "could not read %d loads from %s\n", 3 * n_nodes, a["--load-file"].c_str());
            return 2;
        }
    } else if (a.count("--node-load")) {
        int node;
        double fx;
        std::sscanf(a["--node-load"].c_str(), "%d,%lf", &node, &fx);
        f(3 * node + 0) = fx;
    } else {
        std::vector<int> tip;
        for (int i = 0; i < n_nodes; i++)
            if (std::abs(hx(i) - xmax) < 5e-4 * L) tip.push_back(i);
        const Real load = num(a, "--load", -1000.0);
        for (int i : tip) f(3 * i + 2) = load / (Real)tip.size();
    }
    b->SetNodalFixed(to_vectorxi(fixed));
    b->SetExternalForce(f);
    if (t10)
        b->Setup(Quadrature::tet5pt_x, Quadrature::tet5pt_y, Quadrature::tet5pt_z, Quadrature::tet5pt_weights, hx, hy,
                 hz, elems);
    else
        b->Setup(Quadrature::tet1pt_x, Quadrature::tet1pt_y, Quadrature::tet1pt_z, Quadrature::tet1pt_weights, hx, hy,
                 hz, elems);
    b->SetDensity(num(a, "--rho0", 7850.0));
    b->SetDamping(num(a, "--eta", 0.0), num(a, "--lambda-d", 0.0));
    b->SetSVK(num(a, "--E", 2.1e11), num(a, "--nu", 0.3));
    if (a.count("--mr")) {  // --mr MU10,MU01,KAPPA: compressible Mooney-Rivlin instead of SVK
        double mu10, mu01, kappa;
        if (std::sscanf(a["--mr"].c_str(), "%lf,%lf,%lf", &mu10, &mu01, &kappa) != 3) {
            std::fprintf(stderr, "--mr expects MU10,MU01,KAPPA\n");
            return 2;
        }
        b->SetMooneyRivlin(mu10, mu01, kappa);
    }
    b->CalcDnDuPre();
    if (a.count("--start-file")) {
        VectorXR sx(n_nodes), sy(n_nodes), sz(n_nodes);
        std::ifstream sf(a["--start-file"], std::ios::binary);
        for (VectorXR* v : {&sx, &sy, &sz}) sf.read(reinterpret_cast<char*>(v->data()), sizeof(double) * n_nodes);
        if (!sf) {
            std::fprintf(stderr, "could not read %d positions from %s\n", 3 * n_nodes, a["--start-file"].c_str());
            return 2;
        }
        b->UpdatePositions(sx, sy, sz);
    }
    auto t0 = std::chrono::steady_clock::now();
    std::string extra;
    if (mode == "linear_static") {
        LinearStaticSolver<TData> solver(b, num(a, "--tol", 1e-10), (int)num(a, "--max-iter", 50000));
        solver.UseAssembledOperator(num(a, "--assembled", 0) != 0);
        // --two-level DEG: the FP32 two-level preconditioned CG with DEG Chebyshev steps
        if (num(a, "--two-level", 0) > 0) solver.UseTwoLevelPreconditioner(true, (int)num(a, "--two-level", 0));
        fea.AddSolver("linear_static", &solver);
        // --solves N: N Solve() calls in a row (each re-linearises at the moved positions)
        const int n_solves = std::max(1, (int)num(a, "--solves", 1));
        std::string each;
        for (int k = 0; k < n_solves; k++) {
            fea.GetSolver("linear_static")->Solve();
            each += std::string(k ? "," : "") + (solver.LastSolveWasMatrixFree() ? "true" : "false");
        }
        extra = ", \"iters\": " + std::to_string(solver.GetLastIterCount()) +
                ", \"residual\": " + sci(solver.GetLastResidual()) +
                ", \"matrix_free\": " + (solver.LastSolveWasMatrixFree() ? "true" : "false") +
                ", \"two_level\": " + (solver.LastSolveWasTwoLevel() ? "true" : "false") +
                ", \"matrix_free_each\": [" + each + "]";
    } else {
        b->CalcMassMatrix();
        b->CalcConstraintData();
        b->ConvertToCSR_ConstraintJacT();
        if (a.count("--jac-out")) {
            b->ConvertToCSR_ConstraintJac();
            write_jacobians(a["--jac-out"], b);
        }
        const int steps = (int)num(a, "--steps", 3);
        // --update-fixed-after K: the fixed set changes before step K (prescribed motion
        // switching supports), as FEAT10Data.cu:674-751 allows between Solve calls
        const int upd_after = (int)num(a, "--update-fixed-after", -1);
        size_t n_fixed_now = fixed.size();
        auto maybe_update = [&](int k) {
            if (k != upd_after) return;
            const std::vector<int> f2 = select_nodes(a.count("--fix2") ? a["--fix2"] : "x0", hx, hy, hz);
            b->UpdateNodalFixed(to_vectorxi(f2));
            b->CalcConstraintData();
            b->ConvertToCSR_ConstraintJacT();
            n_fixed_now = f2.size();
        };
        std::string iters;
        if (mode == "nesterov") {
            SyncedNesterovSolver solver(b, b->get_n_constraint());
            solver.Setup();
            SyncedNesterovParams p = {num(a, "--alpha", 1e-8), num(a, "--rho", 1e14), num(a, "--inner-tol", 1e-6),
                                      num(a, "--outer-tol", 1e-6), (int)num(a, "--max-outer", 5),
                                      (int)num(a, "--max-inner", 300), num(a, "--dt", 1e-3)};
            solver.SetParameters(&p);
            fea.AddSolver("nesterov", &solver);
            for (int k = 0; k < steps; k++) {
                maybe_update(k);
                fea.GetSolver("nesterov")->Solve();
                iters += (k ? "," : "") + std::to_string(solver.GetLastInnerIterations());
            }
        } else {
            SyncedAdamWParams p = {num(a, "--lr", 2e-4), 0.9, 0.999, 1e-8, num(a, "--wd", 1e-4),
                                   num(a, "--lr-decay", 0.995), num(a, "--inner-tol", 1e-1),
                                   num(a, "--outer-tol", 1e-6), num(a, "--rho", 1e14), (int)num(a, "--max-outer", 5),
                                   (int)num(a, "--max-inner", 300), num(a, "--dt", 1e-3), 10, 0.0};
            if (mode == "adamw") {
                SyncedAdamWSolver solver(b, b->get_n_constraint());
                solver.Setup();
                solver.SetParameters(&p);
                fea.AddSolver("adamw", &solver);
                for (int k = 0; k < steps; k++) {
                    maybe_update(k);
                fea.GetSolver("adamw")->Solve();
                    iters += (k ? "," : "") + std::to_string(solver.GetLastInnerIterations());
                }
            } else {
                SyncedAdamWNocoopSolver solver(b, b->get_n_constraint());
                solver.Setup();
                solver.SetParameters(&p);
                fea.AddSolver("adamw_nocoop", &solver);
                for (int k = 0; k < steps; k++) {
                    maybe_update(k);
                fea.GetSolver("adamw_nocoop")->Solve();
                    iters += (k ? "," : "") + std::to_string(solver.GetLastInnerIterations());
                }
            }
        }
        extra = ", \"inner_iters\": [" + iters + "], \"n_fixed_final\": " + std::to_string(n_fixed_now);
    }
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    VectorXR x, y, z;
    b->RetrievePositionToCPU(x, y, z);
    if (a.count("--out")) write_positions(a["--out"], x, y, z);
    if (a.count("--vtk")) b->WriteOutputVTK(a["--vtk"]);
    if (a.count("--vtk-feat")) {
        const bool ok = t10 ? ANCFCPUUtils::WriteFEAT10ToVTK(a["--vtk-feat"], nodes, elems, x, y, z)
                            : ANCFCPUUtils::WriteFEAT4ToVTK(a["--vtk-feat"], nodes, elems, x, y, z);
        if (!ok) return 1;
    }
    Real tip = 0.0;
    for (int i = 0; i < n_nodes; i++)
        if (std::abs(hx(i) - xmax) < 5e-4 * L) tip = std::min(tip, z(i) - hz(i));
    std::printf("{\"mode\": \"%s\", \"n_elem\": %d, \"n_nodes\": %d, \"n_fixed\": %zu, \"seconds\": %.6f, "
                "\"tip_uz\": %.12e%s}\n",
                mode.c_str(), n_el, n_nodes, fixed.size(), secs, tip, extra.c_str());
    b->Destroy();
    return 0;
}
}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: tlfea_driver linear_static|nesterov|adamw|adamw_nocoop --mesh SPEC [...]\n");
        return 2;
    }
    const std::string mode = argv[1];
    auto a = parse(argc, argv);
    MatrixXR nodes;
    MatrixXi elems;
    if (!a.count("--mesh") || !load_mesh(a["--mesh"], nodes, elems)) {
        std::fprintf(stderr, "could not load mesh\n");
        return 2;
    }
    if (mode == "write_vtk") {  // no device work: the reference's VTK writers on a sheared mesh
        const Real sh = num(a, "--shear", 0.01);
        const std::ptrdiff_t n = nodes.rows();
        VectorXR x(n), y(n), z(n);
        for (std::ptrdiff_t i = 0; i < n; i++) {
            const Real X = nodes(i, 0), Y = nodes(i, 1), Z = nodes(i, 2);
            x(i) = X + sh * Y * Z;
            y(i) = Y - sh * X * Z;
            z(i) = Z + sh * X * Y;
        }
        const bool ok = elems.cols() == 10 ? ANCFCPUUtils::WriteFEAT10ToVTK(a["--vtk-feat"], nodes, elems, x, y, z)
                                           : ANCFCPUUtils::WriteFEAT4ToVTK(a["--vtk-feat"], nodes, elems, x, y, z);
        std::printf("{\"n_nodes\": %td, \"n_elem\": %td, \"ok\": %s}\n", nodes.rows(), elems.rows(),
                    ok ? "true" : "false");
        return ok ? 0 : 1;
    }
    if (mode == "mesh_info") {  // no device work: dumps the mesh as raw float64 nodes + int32 elements
        if (a.count("--out")) {
            std::ofstream f(a["--out"], std::ios::binary);
            f.write(reinterpret_cast<const char*>(nodes.data()), sizeof(double) * nodes.size());
            f.write(reinterpret_cast<const char*>(elems.data()), sizeof(int) * elems.size());
        }
        std::printf("{\"n_nodes\": %td, \"n_elem\": %td, \"nn\": %td}\n", nodes.rows(), elems.rows(), elems.cols());
        return 0;
    }
    try {
        if (elems.cols() == 10) return run<GPU_FEAT10_Data>(mode, a, nodes, elems);
        return run<GPU_FEAT4_Data>(mode, a, nodes, elems);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
