// dmri-solve: command-line driver over the libdmri C ABI.
//
// Accepts the options of the reference drivers and prints their result lines:
// Meshes: DOLFIN XML, or ASCII gmsh 2.x (.msh/.gmsh) read as dolfin-convert would convert it.
//   GCloud style (GCloudDmriSolver.py:61-133):  -f mesh.xml -M 0|1 -b B -q Q -p kappa -d delta -D Delta
//                                               -k dt -K D0 -T2 T2 -gdir gx gy gz -pdir px py pz
//                                               -IsPeriodic 0|1 (-pdir: pseudo-periodic faces normal to
//                                               the axes with p != 0, weakly imposed, or strongly with
//                                               -IsPeriodic 1 on a periodic mesh)
//   comri style (comri/two-comp/hpc-fenics-cpp/main.cpp:581-648): -m mesh.xml -N steps -b B -q G -d delta
//                                               -D Delta -k dt -K D0 -p kappa -v gx gy gz -j nskip
// Extra: -phase file (per-cell phase, one integer per line, or a DOLFIN mesh_function /
// mesh_value_collection file whose values are partition markers taken modulo 2, as
// CreatePhaseFunc does, DmriFemLib.py:758-765), -variant gcloud|comri|comri-until-T,
// -s 1 -j nskip: write u after every step whose 0-based counter is a multiple of nskip as legacy
// VTK files dir/u000000.vtk, ... (comri's u.bin snapshots, main.cpp:851,886-890; dir = -o, or -f
// when the mesh comes with -m as in comri, default "results").
// -theta, -rtol, -atol, -maxit, -device, -convert out.xml (write the mesh as DOLFIN XML and stop,
// as dolfin-convert mesh.msh mesh.xml does; no device is used).
//
// Sequences: gcloud = loop t < T + k, t += k, explicit side lagged one step, PGSE profile
// exclusive at T (DmriFemLib.py:897-910, GCloudDmriSolver.py:188-193); comri = same f on
// both sides, loop t < T + dt, FT inclusive at Delta + delta (main.cpp:87-92,296-318);
// comri-until-T = the multilayer loop t < T (comri/multilayer/hpc-fenics-cpp/main.cpp:301).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include <sys/stat.h>

#include "dmri.h"

namespace {

struct Mesh {
    std::vector<double> xyz;
    std::vector<int32_t> tets;
};

bool attr(const std::string &line, const char *name, std::string &out) {
    // a whitespace before the name, so that x=" does not match inside index="
    std::string key = std::string(" ") + name + "=\"";
    size_t p = line.find(key);
    if (p == std::string::npos) return false;
    p += key.size();
    size_t e = line.find('"', p);
    if (e == std::string::npos) return false;
    out = line.substr(p, e - p);
    return true;
}

// Vertices and cells are placed by their index attribute; an index must lie below the number of
// entries of its kind (as the Python reader, which sizes the arrays by that count).  A malformed
// number makes the file unreadable (false), never an exception.
bool read_dolfin_xml(const std::string &path, Mesh &m) {
    std::ifstream in(path);
    if (!in) return false;
    std::string line;
    std::vector<std::pair<size_t, std::array<double, 3>>> verts;
    std::vector<std::pair<size_t, std::array<int32_t, 4>>> cells;
    try {
        while (std::getline(in, line)) {
            if (line.find("<vertex ") != std::string::npos) {
                std::string si, sx, sy, sz;
                if (!attr(line, "index", si) || !attr(line, "x", sx) || !attr(line, "y", sy) || !attr(line, "z", sz))
                    return false;
                verts.push_back({std::stoul(si), {std::stod(sx), std::stod(sy), std::stod(sz)}});
            } else if (line.find("<tetrahedron ") != std::string::npos) {
                std::string si, a, b, c, d;
                if (!attr(line, "index", si) || !attr(line, "v0", a) || !attr(line, "v1", b) ||
                    !attr(line, "v2", c) || !attr(line, "v3", d))
                    return false;
                cells.push_back({std::stoul(si), {std::stoi(a), std::stoi(b), std::stoi(c), std::stoi(d)}});
            }
        }
    } catch (const std::exception &) {  // std::invalid_argument / std::out_of_range from stoul, stod, stoi
        return false;
    }
    if (verts.empty() || cells.empty()) return false;
    m.xyz.assign(3 * verts.size(), 0.0);
    m.tets.assign(4 * cells.size(), 0);
    for (const auto &v : verts) {
        if (v.first >= verts.size()) return false;
        for (int d = 0; d < 3; ++d) m.xyz[3 * v.first + d] = v.second[d];
    }
    for (const auto &c : cells) {
        if (c.first >= cells.size()) return false;
        for (int k = 0; k < 4; ++k) m.tets[4 * c.first + k] = c.second[k];
    }
    return true;
}

// ASCII gmsh 2.x, read as dolfin-convert's gmsh2xml turns it into the DOLFIN XML mesh the
// reference solvers load (comri/meshes/neuron_download/dolfin-convert.py:357-597): the tetrahedra
// (element type 4) in file order are the cells; the nodes they use are the vertices, numbered in
// $Nodes order; each coordinate passes through the converter's "%g" (write_vertex, 889-892).
// The four nodes after an element's tags are taken for any tag count (the converter needs two).
bool read_gmsh(const std::string &path, Mesh &m) {
    std::ifstream in(path);
    if (!in) return false;
    std::string line;
    auto next_line = [&]() -> bool { return (bool)std::getline(in, line); };
    auto trim = [](std::string &s) {
        while (!s.empty() && (s.back() == '\r' || s.back() == ' ')) s.pop_back();
    };
    std::vector<long> node_id;
    std::vector<double> xyz;
    std::vector<long> tet_nodes;
    bool fmt_ok = false, have_nodes = false;
    while (next_line()) {
        trim(line);
        if (line == "$MeshFormat") {
            if (!next_line()) return false;
            std::istringstream ss(line);
            std::string ver;
            int ftype = -1;
            ss >> ver >> ftype;
            fmt_ok = !ver.empty() && ver[0] == '2' && ftype == 0;
            if (!fmt_ok) return false;
        } else if (line == "$Nodes") {
            long n = 0;
            if (!next_line() || (n = std::atol(line.c_str())) <= 0) return false;
            char buf[64];
            for (long i = 0; i < n; ++i) {  // grown line by line: a count past the end fails at EOF
                if (!next_line()) return false;
                std::istringstream ss(line);
                long id;
                double c[3];
                if (!(ss >> id >> c[0] >> c[1] >> c[2]) || id < 0) return false;
                node_id.push_back(id);
                for (int d = 0; d < 3; ++d) {
                    std::snprintf(buf, sizeof(buf), "%g", c[d]);
                    xyz.push_back(std::strtod(buf, nullptr));
                }
            }
            have_nodes = true;
        } else if (line == "$Elements") {
            long n = 0;
            if (!next_line() || (n = std::atol(line.c_str())) <= 0) return false;
            for (long i = 0; i < n; ++i) {
                if (!next_line()) return false;
                std::istringstream ss(line);
                long id, type, ntags, t;
                if (!(ss >> id >> type >> ntags)) return false;
                for (long k = 0; k < ntags; ++k)
                    if (!(ss >> t)) return false;
                if (type != 4) continue;
                for (int k = 0; k < 4; ++k) {
                    if (!(ss >> t)) return false;
                    tet_nodes.push_back(t);
                }
            }
        }
    }
    if (!fmt_ok || !have_nodes || tet_nodes.empty()) return false;
    // node id -> position in $Nodes: a table over the ids when they are dense (the usual 1..n),
    // a hash map when they are sparse
    long maxid = 0;
    for (long v : node_id) maxid = std::max(maxid, v);
    const bool dense = maxid <= 4 * (long)node_id.size() + 1024;
    std::vector<long> table(dense ? maxid + 1 : 0, -1);
    std::unordered_map<long, long> hashed;
    for (size_t i = 0; i < node_id.size(); ++i) {
        if (dense) table[node_id[i]] = (long)i;
        else hashed[node_id[i]] = (long)i;
    }
    auto where = [&](long id) -> long {
        if (dense) return id >= 0 && id <= maxid ? table[id] : -1;
        auto it = hashed.find(id);
        return it == hashed.end() ? -1 : it->second;
    };
    std::vector<long> pos_of(tet_nodes.size());
    std::vector<char> used(node_id.size(), 0);
    for (size_t k = 0; k < tet_nodes.size(); ++k) {
        const long p = where(tet_nodes[k]);
        if (p < 0) return false;
        pos_of[k] = p;
        used[p] = 1;
    }
    std::vector<int32_t> renum(node_id.size(), -1);
    int32_t nv = 0;
    for (size_t i = 0; i < node_id.size(); ++i)
        if (used[i]) {
            renum[i] = nv++;
            m.xyz.insert(m.xyz.end(), xyz.begin() + 3 * i, xyz.begin() + 3 * i + 3);
        }
    m.tets.resize(tet_nodes.size());
    for (size_t k = 0; k < tet_nodes.size(); ++k) m.tets[k] = renum[pos_of[k]];
    return true;
}

// DOLFIN XML with every coordinate printed to 17 significant digits (exact round trip)
bool write_dolfin_xml(const std::string &path, const Mesh &m) {
    FILE *f = std::fopen(path.c_str(), "w");
    if (!f) return false;
    const size_t nv = m.xyz.size() / 3, nc = m.tets.size() / 4;
    std::fprintf(f, "<?xml version=\"1.0\"?>\n<dolfin xmlns:dolfin=\"http://fenicsproject.org\">\n"
                    "  <mesh celltype=\"tetrahedron\" dim=\"3\">\n    <vertices size=\"%zu\">\n", nv);
    for (size_t i = 0; i < nv; ++i)
        std::fprintf(f, "      <vertex index=\"%zu\" x=\"%.17g\" y=\"%.17g\" z=\"%.17g\" />\n", i, m.xyz[3 * i],
                     m.xyz[3 * i + 1], m.xyz[3 * i + 2]);
    std::fprintf(f, "    </vertices>\n    <cells size=\"%zu\">\n", nc);
    for (size_t i = 0; i < nc; ++i)
        std::fprintf(f, "      <tetrahedron index=\"%zu\" v0=\"%d\" v1=\"%d\" v2=\"%d\" v3=\"%d\" />\n", i,
                     m.tets[4 * i], m.tets[4 * i + 1], m.tets[4 * i + 2], m.tets[4 * i + 3]);
    std::fprintf(f, "    </cells>\n  </mesh>\n</dolfin>\n");
    return std::fclose(f) == 0;
}

// Legacy VTK unstructured grid (binary, big-endian) of the state u after one time step: the mesh
// and, per compartment c, the point fields u<c>_re and u<c>_im (0 where c has no DOF).
void put_be(FILE *f, const void *p, size_t bytes, size_t count) {
    std::vector<unsigned char> b(bytes * count);
    const unsigned char *s = static_cast<const unsigned char *>(p);
    for (size_t i = 0; i < count; ++i)
        for (size_t k = 0; k < bytes; ++k) b[i * bytes + k] = s[i * bytes + bytes - 1 - k];
    std::fwrite(b.data(), 1, b.size(), f);
}

bool write_vtk(const std::string &path, const Mesh &m, int ncomp, const std::vector<int32_t> &dof_of,
               const double *u, int step) {
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const size_t nv = m.xyz.size() / 3, nc = m.tets.size() / 4;
    std::fprintf(f, "# vtk DataFile Version 3.0\ndmri-solve u after time step %d\nBINARY\n"
                    "DATASET UNSTRUCTURED_GRID\nPOINTS %zu double\n", step, nv);
    put_be(f, m.xyz.data(), 8, 3 * nv);
    std::fprintf(f, "\nCELLS %zu %zu\n", nc, 5 * nc);
    std::vector<int32_t> cells(5 * nc);
    for (size_t i = 0; i < nc; ++i) {
        cells[5 * i] = 4;
        for (int k = 0; k < 4; ++k) cells[5 * i + 1 + k] = m.tets[4 * i + k];
    }
    put_be(f, cells.data(), 4, cells.size());
    std::fprintf(f, "\nCELL_TYPES %zu\n", nc);
    std::vector<int32_t> types(nc, 10);  // VTK_TETRA
    put_be(f, types.data(), 4, nc);
    std::fprintf(f, "\nPOINT_DATA %zu\n", nv);
    std::vector<double> field(nv);
    for (int c = 0; c < ncomp; ++c)
        for (int part = 0; part < 2; ++part) {
            for (size_t v = 0; v < nv; ++v) {
                const int32_t d = dof_of[(size_t)c * nv + v];
                field[v] = d < 0 ? 0.0 : u[2 * (size_t)d + part];
            }
            std::fprintf(f, "SCALARS u%d_%s double 1\nLOOKUP_TABLE default\n", c, part ? "im" : "re");
            put_be(f, field.data(), 8, nv);
            std::fprintf(f, "\n");
        }
    return std::fclose(f) == 0;
}

bool ends_with(const std::string &s, const char *suffix) {
    const size_t n = std::strlen(suffix);
    return s.size() >= n && s.compare(s.size() - n, n, suffix) == 0;
}

// per-cell phase: plain integers, or <value cell_index="i" ... value="m" /> entries (marker % 2)
bool read_phase(const std::string &path, size_t nc, std::vector<int32_t> &phase) {
    std::ifstream in(path);
    if (!in) return false;
    phase.assign(nc, 0);
    std::string line;
    size_t plain = 0;
    bool xml = false;
    try {
        while (std::getline(in, line)) {
            if (line.find("<value ") != std::string::npos) {
                xml = true;
                std::string ci, val;
                if (!attr(line, "cell_index", ci) || !attr(line, "value", val)) return false;
                size_t i = std::stoul(ci);
                if (i >= nc) return false;
                phase[i] = (std::stoi(val) % 2 + 2) % 2;  // marker mod 2, non-negative
            } else if (!xml) {
                std::istringstream ss(line);
                long v;
                while (ss >> v) {
                    if (plain >= nc) return false;
                    phase[plain++] = (int32_t)((v % 2 + 2) % 2);
                }
            }
        }
    } catch (const std::exception &) {  // a malformed cell_index or value
        return false;
    }
    return xml || plain == nc;
}

void usage() {
    std::fprintf(stderr,
                 "usage: dmri-solve (-f|-m) mesh.xml|mesh.msh [-M 0|1] [-phase file] [-b B | -q Q] [-d delta] [-D Delta]\n"
                 "                  [-k dt | -N steps] [-K D0] [-p kappa] [-T2 T2] [-gdir|-v gx gy gz]\n"
                 "                  [-pdir px py pz] [-IsPeriodic 0|1] [-r 0]\n"
                 "                  [-variant gcloud|comri|comri-until-T] [-theta t] [-rtol r] [-atol a]\n"
                 "                  [-maxit n] [-device d] [-convert out.xml] [-s 0|1 [-j nskip] [-o dir]]\n");
}

}  // namespace

int main(int argc, char **argv) {
    // defaults of GCloudDmriSolver.py:52-55
    std::string fmesh, fmesh_m, farg, fphase, fconvert, dir = "results", variant = "gcloud";
    int nskip = 5, is_save = 0;
    int multiple = 0, nsteps = 100, maxit = 100000, device = 0;
    double b = 1000, qin = -1, kappa = 1e-5, delta = 10600, Delta = 43100, dt = 200, D0 = 3e-3, T2 = 1e16;
    double theta = 0.5, rtol = 1e-9, atol = 1e-10;
    double g[3] = {0, 1, 0};
    int32_t pdir[3] = {0, 0, 0};  // -pdir (GCloudDmriSolver.py:126-131)
    int is_periodic = 0, nrefine = 0;
    bool have_dt = false, have_b = false, have_q = false;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&](int k = 1) -> const char * {
            if (i + k >= argc) {
                usage();
                std::exit(2);
            }
            return argv[i + k];
        };
        if (a == "-f") { farg = next(); ++i; }
        else if (a == "-m") { fmesh_m = next(); ++i; }
        else if (a == "-j") { nskip = std::atoi(next()); ++i; }
        else if (a == "-s") { is_save = std::atoi(next()); ++i; }
        else if (a == "-o") { dir = next(); ++i; }
        else if (a == "-M") { multiple = std::atoi(next()); ++i; }
        else if (a == "-phase" || a == "-c") { fphase = next(); ++i; multiple = 1; }
        else if (a == "-b") { b = std::atof(next()); have_b = true; ++i; }
        else if (a == "-q") { qin = std::atof(next()); have_q = true; ++i; }
        else if (a == "-p") { kappa = std::atof(next()); ++i; }
        else if (a == "-d") { delta = std::atof(next()); ++i; }
        else if (a == "-D") { Delta = std::atof(next()); ++i; }
        else if (a == "-k") { dt = std::atof(next()); have_dt = true; ++i; }
        else if (a == "-N") { nsteps = std::atoi(next()); ++i; }
        else if (a == "-K") { D0 = std::atof(next()); ++i; }
        else if (a == "-T2") { T2 = std::atof(next()); ++i; }
        else if (a == "-gdir" || a == "-v") {
            g[0] = std::atof(next(1)); g[1] = std::atof(next(2)); g[2] = std::atof(next(3)); i += 3;
        }
        else if (a == "-variant") { variant = next(); ++i; }
        else if (a == "-theta") { theta = std::atof(next()); ++i; }
        else if (a == "-rtol") { rtol = std::atof(next()); ++i; }
        else if (a == "-atol") { atol = std::atof(next()); ++i; }
        else if (a == "-maxit") { maxit = std::atoi(next()); ++i; }
        else if (a == "-device") { device = std::atoi(next()); ++i; }
        else if (a == "-convert") { fconvert = next(); ++i; }
        else if (a == "-pdir") {
            pdir[0] = std::atoi(next(1)); pdir[1] = std::atoi(next(2)); pdir[2] = std::atoi(next(3)); i += 3;
        }
        else if (a == "-IsPeriodic" || a == "-isperiodic") { is_periodic = std::atoi(next()); ++i; }
        else if (a == "-r") { nrefine = std::atoi(next()); ++i; }
        else if (a == "-h" || a == "--help") { usage(); return 0; }
        else { std::fprintf(stderr, "unknown option %s\n", a.c_str()); usage(); return 2; }
    }
    // comri: -m mesh -f output directory (main.cpp:599-615); GCloud: -f mesh
    if (!fmesh_m.empty()) {
        fmesh = fmesh_m;
        if (!farg.empty()) dir = farg;
    } else {
        fmesh = farg;
    }
    if (fmesh.empty()) { usage(); return 2; }
    if (is_save && nskip <= 0) {
        std::fprintf(stderr, "-j nskip must be positive\n");
        return 2;
    }
    if (nrefine != 0) {
        // comri's -r refines the mesh uniformly before the solve (two-comp main.cpp:615); meshes are
        // taken as given here
        std::fprintf(stderr, "-r %d: mesh refinement is not supported (refine the mesh beforehand)\n", nrefine);
        return 2;
    }
    // -IsPeriodic 1 with -pdir: the strong pseudo-periodic form (PeriodicBD, DmriFemLib.py:327-375)
    const bool periodic = pdir[0] || pdir[1] || pdir[2];
    const bool comri = variant != "gcloud";
    if (variant != "gcloud" && variant != "comri" && variant != "comri-until-T") {
        std::fprintf(stderr, "unknown variant %s\n", variant.c_str());
        return 2;
    }
    const double gnorm_in = std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
    if (!(gnorm_in > 1e-10)) {
        std::printf("|g|=0! Please check again the gradient directions!\n");  // DmriFemLib.py:823
        return 1;
    }
    const double T = Delta + delta;
    const double int4gb = delta * delta * (Delta - delta / 3.0);
    if (!(int4gb > 0)) {
        std::fprintf(stderr, "Delta - delta/3 must be positive\n");
        return 1;
    }
    // q = gamma G from b (DmriFemLib.py:840-842, main.cpp:187-191) or from -q (G in T/m for comri)
    double q;
    const double g_ratio = 2.675e8;
    if (have_q && !have_b) {
        q = comri ? qin * g_ratio * 1e-12 : qin;
        b = q * q * int4gb;
    } else {
        q = std::sqrt(b) / std::sqrt(int4gb);
    }
    if (comri && !have_dt) dt = T / nsteps;

    // time profile per step, and F = int_0^t f at the explicit side's time (the weak periodic terms'
    // phase: itime_profile(tp), DmriFemLib.py:901-902)
    std::vector<double> fA, fB, FA, FB;
    auto f_py = [&](double s) { return s < delta ? 1.0 : (s < Delta ? 0.0 : (s < T ? -1.0 : 0.0)); };
    auto F_py = [&](double s) { return s < delta ? s : (s < Delta ? delta : (s < T ? delta - (s - Delta) : 0.0)); };
    auto f_c = [&](double t) { return 1.0 * (t >= 0 && t < delta) - 1.0 * (t >= Delta && t <= Delta + delta); };
    auto F_c = [&](double t) {
        return t < 0 ? 0.0 : (t < delta ? t : (t < Delta ? delta : (t <= Delta + delta ? delta - (t - Delta) : 0.0)));
    };
    if (!comri) {
        double t = 0, tp = 0;
        while (t < T + dt) {
            fA.push_back(f_py(t));
            fB.push_back(f_py(tp));
            FA.push_back(F_py(t));
            FB.push_back(F_py(tp));
            tp = t;
            t += dt;
        }
    } else {
        double t = 0;
        const double end = variant == "comri-until-T" ? T : T + dt;
        while (t < end) {
            fA.push_back(f_c(t));
            FA.push_back(F_c(t));
            FB.push_back(F_c(t));
            t += dt;
        }
        fB = fA;
    }

    Mesh m;
    const bool gmsh = ends_with(fmesh, ".msh") || ends_with(fmesh, ".gmsh");
    if (gmsh ? !read_gmsh(fmesh, m) : !read_dolfin_xml(fmesh, m)) {
        std::fprintf(stderr, "cannot read %s tetrahedral mesh %s\n", gmsh ? "ASCII gmsh 2.x" : "DOLFIN XML",
                     fmesh.c_str());
        return 1;
    }
    if (!fconvert.empty()) {
        // mesh conversion only (dolfin-convert mesh.msh mesh.xml); no device is touched
        if (!write_dolfin_xml(fconvert, m)) {
            std::fprintf(stderr, "cannot write %s\n", fconvert.c_str());
            return 1;
        }
        std::printf("converted %s: %zu vertices, %zu tetrahedra -> %s\n", fmesh.c_str(), m.xyz.size() / 3,
                    m.tets.size() / 4, fconvert.c_str());
        return 0;
    }
    const int64_t nv = (int64_t)m.xyz.size() / 3, nc = (int64_t)m.tets.size() / 4;
    std::vector<int32_t> phase;
    if (multiple) {
        if (fphase.empty() || !read_phase(fphase, (size_t)nc, phase)) {
            std::fprintf(stderr, "two compartments need -phase with one value per cell\n");
            return 1;
        }
    }
    dmri_ctx *ctx = nullptr;
    auto check = [&](int rc, const char *what) {
        if (rc != DMRI_OK) {
            std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, dmri_last_error());
            if (ctx) dmri_destroy(ctx);
            std::exit(1);
        }
    };
    check(dmri_create(&ctx, device), "dmri_create");
    check(dmri_mesh_set(ctx, nv, m.xyz.data(), nc, m.tets.data(), multiple ? phase.data() : nullptr),
          "dmri_mesh_set");
    dmri_coeffs co{};
    co.D[0] = co.D[1] = D0;
    co.invT2[0] = co.invT2[1] = 1.0 / T2;
    co.kappa = kappa;
    check(dmri_coeffs_set(ctx, &co), "dmri_coeffs_set");
    if (periodic && is_periodic) check(dmri_periodic_strong_set(ctx, pdir), "dmri_periodic_strong_set");
    else if (periodic) check(dmri_periodic_set(ctx, pdir), "dmri_periodic_set");
    check(dmri_assemble(ctx), "dmri_assemble");
    dmri_seq seq{};
    seq.nsteps = (int32_t)fA.size();
    seq.f_implicit = fA.data();
    seq.f_explicit = fB.data();
    seq.dt = dt;
    seq.theta = theta;
    seq.q = q;
    seq.gdir[0] = g[0];
    seq.gdir[1] = g[1];
    seq.gdir[2] = g[2];
    if (periodic) {
        seq.F_explicit = FB.data();
        seq.F_implicit = FA.data();
    }
    dmri_info info0{};
    check(dmri_get_info(ctx, &info0), "dmri_get_info");
    std::vector<double> snaps;
    if (is_save) {
        const size_t nsnap = ((size_t)seq.nsteps + nskip - 1) / nskip;
        snaps.assign(nsnap * 2 * (size_t)info0.ndof, 0.0);
        seq.snap_every = nskip;
        seq.snap_out = snaps.data();
    }
    dmri_solver_opts op{};
    op.rtol = rtol;
    op.atol = atol;
    op.maxit = maxit;
    double sig[2];
    dmri_stats st{};
    check(dmri_solve(ctx, &seq, &op, sig, &st), "dmri_solve");
    if (is_save) {
        // u after every step with step_counter % nskip == 0, numbered by write as File does
        // (main.cpp:851,886-890): dir/u000000.vtk, dir/u000001.vtk, ...
        ::mkdir(dir.c_str(), 0755);
        const int ncomp = (int)info0.ncomp;
        std::vector<int32_t> dof_of((size_t)ncomp * nv);
        std::vector<int64_t> rp(info0.ndof + 1);
        std::vector<int32_t> cl(info0.nnz);
        check(dmri_export_pattern(ctx, rp.data(), cl.data(), dof_of.data()), "dmri_export_pattern");
        const size_t nsnap = snaps.size() / (2 * (size_t)info0.ndof);
        for (size_t j = 0; j < nsnap; ++j) {
            char name[32];
            std::snprintf(name, sizeof(name), "/u%06zu.vtk", j);
            if (!write_vtk(dir + name, m, ncomp, dof_of, snaps.data() + j * 2 * (size_t)info0.ndof,
                           (int)(j * nskip))) {
                std::fprintf(stderr, "cannot write %s%s\n", dir.c_str(), name);
                dmri_destroy(ctx);
                return 1;
            }
        }
    }
    dmri_info info{};
    check(dmri_get_info(ctx, &info), "dmri_get_info");
    const double whole = info.volume, elapsed = st.solve_ms * 1e-3;
    if (!comri) {
        // PostProcessing line (DmriFemLib.py:933,972) + log.txt append (982-988)
        char line[1024];
        const double gval = q / g_ratio * 1e12;
        if (multiple)
            std::snprintf(line, sizeof(line),
                          "b: %.3f, g: %.3f, q: %.3e, Signal: %.3e, Normalized signal: %.6e, kappa: %.3e, dt: %.3f, "
                          "hmin: %.3e, hmax: %.3e, whole_vol: %.3f, vol_of_interest: %.3f, elasped time %.3f (s)\n",
                          b, gval, q, sig[0] * whole, sig[0], kappa, dt, info.hmin, info.hmax, whole, whole, elapsed);
        else
            std::snprintf(line, sizeof(line),
                          "b: %.3f, g: %.3f, q: %.3e, Signal: %.3e, Normalized signal: %.6e, dt: %.3f, hmin: %.3e, "
                          "hmax: %.3e, whole_vol: %.3f, vol_of_interest: %.3f, elasped time %.3f (s)\n",
                          b, gval, q, sig[0] * whole, sig[0], dt, info.hmin, info.hmax, whole, whole, elapsed);
        std::fputs(line, stdout);
        if (FILE *lf = std::fopen("log.txt", "a")) {
            std::fputs(line, lf);
            std::fclose(lf);
        }
    } else {
        // comri message (one-comp main.cpp:333; two-comp main.cpp:900): gnorm is q, "q" is G in T/m
        const double G = q / g_ratio * 1e12, gn = gnorm_in;
        if (multiple)
            std::printf("b: %f, gnorm: %e, q: %e, perm: %e, gdir: (%f, %f, %f), s: %f\n", b, q, G, kappa, g[0] / gn,
                        g[1] / gn, g[2] / gn, sig[0]);
        else
            std::printf("b: %f, gnorm: %f, q: %f, gdir: (%f, %f, %f), s: %f\n", b, q, G, g[0] / gn, g[1] / gn,
                        g[2] / gn, sig[0]);
        std::printf("Runtime = %f\n", elapsed);
    }
    std::fprintf(stderr, "[dmri-solve] nv=%lld nc=%lld ndof=%lld nnz=%lld steps=%d krylov=%lld signal=%.12e\n",
                 (long long)nv, (long long)nc, (long long)info.ndof, (long long)info.nnz, seq.nsteps,
                 (long long)st.iters_total, sig[0]);
    dmri_destroy(ctx);
    return 0;
}
