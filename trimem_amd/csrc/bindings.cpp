// bindings.cpp -- pybind11 module trimem_amd.core.
//
// Mirrors the Python surface of bio-phys/trimem's trimem.core
// (src/bindings.cpp:52-1093): the same classes, functions, argument names,
// return kinds and error types, implemented on the C ABI of libtrimem_hip
// (include/trimem_hip.h).  Additions over the reference are keyword-only
// extras: a seed for flip/pflip, flip_injected, and EnergyManager.precision.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <atomic>
#include <chrono>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/trimem_hip.h"

namespace py = pybind11;

namespace {

void check(int rc)
{
    if (rc == TM_OK) return;
    std::string msg = tm_last_error();
    if (rc == TM_E_RANGE) throw py::value_error(msg);
    throw std::runtime_error(msg);
}

// ---------------------------------------------------------------- mesh ----
struct HalfedgeHandle {
    int h = -1;
    HalfedgeHandle() = default;
    explicit HalfedgeHandle(int i) : h(i) {}
    int idx() const { return h; }
    bool is_valid() const { return h >= 0; }
};

struct TriMesh {
    tm_mesh* h = nullptr;
    TriMesh() { check(tm_mesh_create(nullptr, 0, nullptr, 0, &h)); }
    TriMesh(const double* p, int n, const int* f, int m) { check(tm_mesh_create(p, n, f, m, &h)); }
    TriMesh(const TriMesh&) = delete;
    TriMesh& operator=(const TriMesh&) = delete;
    ~TriMesh() { tm_mesh_destroy(h); }
    int nv() const { int a, b, c; check(tm_mesh_counts(h, &a, &b, &c)); return a; }
    int ne() const { int a, b, c; check(tm_mesh_counts(h, &a, &b, &c)); return b; }
    int nf() const { int a, b, c; check(tm_mesh_counts(h, &a, &b, &c)); return c; }
};

// from_points_cells (mesh_py.cpp:12-69): shape checks and empty inputs
std::unique_ptr<TriMesh> from_points_cells(py::array_t<double, py::array::c_style | py::array::forcecast> points,
                                           py::array_t<int, py::array::c_style | py::array::forcecast> cells)
{
    if (points.size() == 0) return std::make_unique<TriMesh>();
    if (points.ndim() != 2 || points.shape(1) != 3) throw std::runtime_error("Array 'points' must have shape (n,3)");
    int n = (int)points.shape(0);
    if (n < 3 || cells.size() == 0) return std::make_unique<TriMesh>(points.data(), n, nullptr, 0);
    if (cells.ndim() != 2 || cells.shape(1) != 3) throw std::runtime_error("Array 'cells' must have shape (n,3)");
    return std::make_unique<TriMesh>(points.data(), n, cells.data(), (int)cells.shape(0));
}

py::array_t<int> fv_indices(TriMesh& m)
{
    int nf = m.nf();
    if (nf == 0) return py::array_t<int>();
    py::array_t<int> out({nf, 3});
    check(tm_mesh_fv_indices(m.h, out.mutable_data()));
    return out;
}

py::array_t<double> points_view(py::object self)
{
    TriMesh& m = self.cast<TriMesh&>();
    size_t n = (size_t)m.nv();
    check(tm_mesh_sync_host(m.h));    // newer device positions, if any
    return py::array_t<double>({n, (size_t)3}, {3 * sizeof(double), sizeof(double)}, tm_mesh_points(m.h), self);
}

// read_mesh (mesh.cpp:7-15 via OpenMesh IO): STL (ASCII and binary, with
// exact-coordinate vertex merging in order of first appearance), OFF and OBJ
// triangle meshes.  A read failure ends the process like the reference.
std::unique_ptr<TriMesh> read_mesh(const std::string& fname)
{
    std::vector<double> pts;
    std::vector<int> faces;
    auto fail = [&]() -> std::unique_ptr<TriMesh> {
        std::fprintf(stderr, "read error on file %s\n", fname.c_str());
        std::exit(1);
    };
    std::ifstream in(fname, std::ios::binary);
    if (!in) return fail();
    std::string ext = fname.size() >= 4 ? fname.substr(fname.size() - 4) : "";
    for (auto& c : ext) c = (char)std::tolower(c);
    if (ext == ".stl") {
        std::vector<char> buf((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
        std::map<std::tuple<float, float, float>, int> vmap;
        auto add = [&](float x, float y, float z) {
            auto k = std::make_tuple(x, y, z);
            auto it = vmap.find(k);
            if (it != vmap.end()) return it->second;
            int id = (int)(pts.size() / 3);
            vmap[k] = id;
            pts.push_back(x); pts.push_back(y); pts.push_back(z);
            return id;
        };
        bool binary = false;
        if (buf.size() >= 84) {
            uint32_t nt;
            std::memcpy(&nt, buf.data() + 80, 4);
            binary = buf.size() == 84 + 50ull * nt;
        }
        if (binary) {
            uint32_t nt;
            std::memcpy(&nt, buf.data() + 80, 4);
            for (uint32_t t = 0; t < nt; t++) {
                const char* r = buf.data() + 84 + 50ull * t + 12;
                float v[9];
                std::memcpy(v, r, sizeof(v));
                int a = add(v[0], v[1], v[2]), b = add(v[3], v[4], v[5]), c = add(v[6], v[7], v[8]);
                if (a != b && b != c && a != c) { faces.push_back(a); faces.push_back(b); faces.push_back(c); }
            }
        } else {
            std::istringstream ss(std::string(buf.begin(), buf.end()));
            std::string tok;
            std::vector<int> tri;
            while (ss >> tok) {
                if (tok == "vertex") {
                    float x, y, z;
                    ss >> x >> y >> z;
                    tri.push_back(add(x, y, z));
                    if (tri.size() == 3) {
                        if (tri[0] != tri[1] && tri[1] != tri[2] && tri[0] != tri[2])
                            faces.insert(faces.end(), tri.begin(), tri.end());
                        tri.clear();
                    }
                }
            }
        }
    } else if (ext == ".off") {
        std::string head;
        in >> head;
        if (head != "OFF") return fail();
        int nv, nf, ne;
        in >> nv >> nf >> ne;
        pts.resize(3 * nv);
        for (auto& v : pts) in >> v;
        for (int f = 0; f < nf; f++) {
            int k;
            in >> k;
            std::vector<int> poly(k);
            for (auto& x : poly) in >> x;
            for (int t = 1; t + 1 < k; t++) { faces.push_back(poly[0]); faces.push_back(poly[t]); faces.push_back(poly[t + 1]); }
        }
        if (!in) return fail();
    } else if (ext == ".obj") {
        std::string line;
        while (std::getline(in, line)) {
            std::istringstream ls(line);
            std::string tag;
            ls >> tag;
            if (tag == "v") {
                double x, y, z;
                ls >> x >> y >> z;
                pts.push_back(x); pts.push_back(y); pts.push_back(z);
            } else if (tag == "f") {
                std::vector<int> poly;
                std::string w;
                while (ls >> w) poly.push_back(std::stoi(w.substr(0, w.find('/'))) - 1);
                for (size_t t = 1; t + 1 < poly.size(); t++) { faces.push_back(poly[0]); faces.push_back(poly[t]); faces.push_back(poly[t + 1]); }
            }
        }
    } else {
        return fail();
    }
    return std::make_unique<TriMesh>(pts.data(), (int)(pts.size() / 3), faces.data(), (int)(faces.size() / 3));
}

// ---------------------------------------------------------- properties ----
struct VertexProperties {
    double area = 0, volume = 0, curvature = 0, bending = 0, tethering = 0, repulsion = 0;
};
VertexProperties vp_of(const tm_props& p)
{
    return VertexProperties{p.area, p.volume, p.curvature, p.bending, p.tethering, p.repulsion};
}

// -------------------------------------------------------------- params ----
enum class BondType : int { Edge = 0, Area = 1, None = 2 };
struct BondParams {
    double lc0 = 1, lc1 = 0, a0 = 1;
    int r = 2;
    BondType type = BondType::Edge;
};
struct SurfaceRepulsionParams {
    double lc1 = 0;
    int r = 2;
    std::string n_search = "cell-list";
    double rlist = 0.1;
    int exclusion_level = 2;
};
struct ContinuationParams {
    double delta = 0, lam = 1;
};
struct EnergyParams {
    double kappa_b = 0, kappa_a = 0, kappa_v = 0, kappa_c = 0, kappa_t = 0, kappa_r = 0;
    double area_frac = 1, volume_frac = 1, curvature_frac = 1;
    BondParams bond_params;
    ContinuationParams continuation_params;
    SurfaceRepulsionParams repulse_params;
};

tm_params to_c(const EnergyParams& e)
{
    tm_params p{};
    p.kappa_b = e.kappa_b; p.kappa_a = e.kappa_a; p.kappa_v = e.kappa_v;
    p.kappa_c = e.kappa_c; p.kappa_t = e.kappa_t; p.kappa_r = e.kappa_r;
    p.area_frac = e.area_frac; p.volume_frac = e.volume_frac; p.curvature_frac = e.curvature_frac;
    p.bond_lc0 = e.bond_params.lc0; p.bond_lc1 = e.bond_params.lc1; p.bond_a0 = e.bond_params.a0;
    p.bond_r = e.bond_params.r; p.bond_type = (int)e.bond_params.type;
    p.rep_lc1 = e.repulse_params.lc1; p.rep_r = e.repulse_params.r;
    const std::string& ns = e.repulse_params.n_search;
    p.rep_nlist = ns == "cell-list" ? 0 : (ns == "verlet-list" ? 1 : -1);
    p.rep_rlist = e.repulse_params.rlist;
    p.rep_exclusion = e.repulse_params.exclusion_level;
    p.cont_delta = e.continuation_params.delta;
    p.cont_lambda = e.continuation_params.lam;
    return p;
}

// --------------------------------------------------------------- nlist ----
struct Point {
    double x, y, z;
};

struct NeighbourList {
    tm_nlist* h = nullptr;
    bool owned = true;
    py::object owner;    // keeps a borrowed list's energy manager alive
    NeighbourList() = default;
    NeighbourList(const NeighbourList&) = delete;
    ~NeighbourList()
    {
        if (owned && h) tm_nlist_destroy(h);
    }

    std::tuple<std::vector<double>, std::vector<int>, std::vector<int>> distance_matrix(TriMesh& m, double dmax)
    {
        long n = 0;
        check(tm_nlist_distance_matrix(h, m.h, dmax, nullptr, nullptr, nullptr, 0, &n));
        std::vector<double> d(n);
        std::vector<int> i(n), j(n);
        if (n) check(tm_nlist_distance_matrix(h, m.h, dmax, d.data(), i.data(), j.data(), n, &n));
        return {d, i, j};
    }

    std::tuple<std::vector<Point>, std::vector<int>> point_distances(TriMesh& m, int pid, double dmax)
    {
        long n = 0;
        check(tm_nlist_point_distances(h, m.h, pid, dmax, nullptr, nullptr, 0, &n));
        std::vector<double> d(3 * n);
        std::vector<int> j(n);
        if (n) check(tm_nlist_point_distances(h, m.h, pid, dmax, d.data(), j.data(), n, &n));
        std::vector<Point> pts(n);
        for (long k = 0; k < n; k++) pts[k] = Point{d[3 * k], d[3 * k + 1], d[3 * k + 2]};
        return {pts, j};
    }
};

// -------------------------------------------------------------- energy ----
struct EnergyManager {
    tm_energy* h = nullptr;
    EnergyManager(TriMesh& m, const EnergyParams& e)
    {
        tm_params p = to_c(e);
        check(tm_energy_create(m.h, &p, &h));
    }
    EnergyManager(const EnergyManager&) = delete;
    ~EnergyManager() { tm_energy_destroy(h); }

    EnergyParams eparams() const
    {
        tm_params p;
        tm_energy_get_params(h, &p);
        EnergyParams e;
        e.kappa_b = p.kappa_b; e.kappa_a = p.kappa_a; e.kappa_v = p.kappa_v;
        e.kappa_c = p.kappa_c; e.kappa_t = p.kappa_t; e.kappa_r = p.kappa_r;
        e.area_frac = p.area_frac; e.volume_frac = p.volume_frac; e.curvature_frac = p.curvature_frac;
        e.bond_params.lc0 = p.bond_lc0; e.bond_params.lc1 = p.bond_lc1; e.bond_params.a0 = p.bond_a0;
        e.bond_params.r = p.bond_r; e.bond_params.type = (BondType)p.bond_type;
        e.repulse_params.lc1 = p.rep_lc1; e.repulse_params.r = p.rep_r;
        e.repulse_params.n_search = p.rep_nlist == 0 ? "cell-list" : "verlet-list";
        e.repulse_params.rlist = p.rep_rlist;
        e.repulse_params.exclusion_level = p.rep_exclusion;
        e.continuation_params.delta = p.cont_delta;
        e.continuation_params.lam = p.cont_lambda;
        return e;
    }
    VertexProperties initial_props() const
    {
        tm_props p;
        tm_energy_initial_props(h, &p);
        return vp_of(p);
    }
    VertexProperties properties(TriMesh& m)
    {
        tm_props p;
        check(tm_energy_properties(h, m.h, &p));
        return vp_of(p);
    }
    double energy(TriMesh& m)
    {
        double e;
        check(tm_energy_energy(h, m.h, &e));
        return e;
    }
    double energy_props(const VertexProperties& v)
    {
        tm_props p{v.area, v.volume, v.curvature, v.bending, v.tethering, v.repulsion};
        double e;
        check(tm_energy_energy_props(h, &p, &e));
        return e;
    }
    py::array_t<double> gradient(TriMesh& m)
    {
        py::array_t<double> out({m.nv(), 3});
        check(tm_energy_gradient(h, m.h, out.mutable_data()));
        return out;
    }
};

// The surrogate force network (tm_surrogate_*).
struct Surrogate {
    tm_surrogate* h = nullptr;
    int nv = 0;
    Surrogate(TriMesh& m, EnergyManager& e, uint64_t seed) : nv(m.nv()) { check(tm_surrogate_create(m.h, e.h, seed, &h)); }
    Surrogate(const Surrogate&) = delete;
    ~Surrogate() { tm_surrogate_destroy(h); }
};

// One HMC step of trimem.mc.hmc (hmc.py:133-158) run on the device; the
// Python samplers in trimem_amd.mc drive it.
struct DeviceHMC {
    tm_hmc* h = nullptr;
    int nv = 0;
    DeviceHMC(TriMesh& m, EnergyManager* e, const std::string& model) : nv(m.nv())
    {
        int md;
        if (model == "energy") md = TM_HMC_ENERGY;
        else if (model == "harmonic") md = TM_HMC_HARMONIC;
        else throw std::runtime_error("Unknown HMC energy model: " + model);
        check(tm_hmc_create(m.h, e ? e->h : nullptr, md, &h));
    }
    DeviceHMC(const DeviceHMC&) = delete;
    ~DeviceHMC() { tm_hmc_destroy(h); }

    // returns the driver of the step: "exact" or "surrogate"
    std::string step(double mass, double time_step, int num_integration_steps, double temperature,
                     py::object momenta, double uniform, uint64_t seed, uint64_t step, double time_step_sq,
                     Surrogate* surrogate, bool guarded)
    {
        tm_hmc_params p{};
        p.mass = mass;
        p.time_step = time_step;
        p.time_step_sq = time_step_sq;
        p.num_integration_steps = num_integration_steps;
        const double* mp = nullptr;
        py::array_t<double, py::array::c_style | py::array::forcecast> a;
        if (!momenta.is_none()) {
            a = py::array_t<double, py::array::c_style | py::array::forcecast>::ensure(momenta);
            if (!a || a.ndim() != 2 || a.shape(0) != nv || a.shape(1) != 3)
                throw std::runtime_error("momenta must have shape (n_vertices, 3)");
            mp = a.data();
        }
        if (surrogate && guarded) {
            int driver = 0;
            check(tm_hmc_step_guarded(h, surrogate->h, &p, temperature, mp, uniform, seed, step, &driver));
            return driver ? "surrogate" : "exact";
        }
        if (surrogate) {
            check(tm_hmc_step_surrogate(h, surrogate->h, &p, temperature, mp, uniform, seed, step));
            return "surrogate";
        }
        check(tm_hmc_step(h, &p, temperature, mp, uniform, seed, step));
        return "exact";
    }
    py::dict result()
    {
        tm_hmc_outcome o;
        check(tm_hmc_result(h, &o));
        py::dict d;
        d["accepted"] = (bool)o.accepted;
        d["dh"] = o.dh;
        d["energy0"] = o.energy0;
        d["energy1"] = o.energy1;
        d["kinetic0"] = o.kinetic0;
        d["kinetic1"] = o.kinetic1;
        d["uniform"] = o.uniform;
        d["n_accepted"] = o.n_accepted;
        d["n_steps"] = o.n_steps;
        return d;
    }
};

py::dict tuning_dict(const tm_tuning& t)
{
    py::dict d;
    d["rep_threads"] = t.rep_threads;
    d["rep_batch"] = t.rep_batch;
    d["pass2_threads"] = t.pass2_threads;
    d["rep_fused"] = t.rep_fused;
    d["pair_cache"] = t.pair_cache;
    d["pflip_window"] = t.pflip_window;
    d["serial_window"] = t.serial_window;
    return d;
}

tm_tuning tuning_of(const py::dict& d)
{
    tm_tuning t{};
    t.rep_threads = d["rep_threads"].cast<int>();
    t.rep_batch = d["rep_batch"].cast<int>();
    t.pass2_threads = d["pass2_threads"].cast<int>();
    t.rep_fused = d.contains("rep_fused") ? d["rep_fused"].cast<int>() : 0;
    t.pair_cache = d.contains("pair_cache") ? d["pair_cache"].cast<int>() : 1;
    t.pflip_window = d.contains("pflip_window") ? d["pflip_window"].cast<int>() : 256;
    t.serial_window = d.contains("serial_window") ? d["serial_window"].cast<int>() : 256;
    return t;
}

// R meshes with their energy managers advanced together (tm_replicas_*).
struct Replicas {
    tm_replicas* h = nullptr;
    std::vector<int> nv;
    Replicas(const std::vector<TriMesh*>& meshes, const std::vector<EnergyManager*>& estores)
    {
        if (meshes.size() != estores.size()) throw std::runtime_error("one energy manager per mesh");
        std::vector<tm_mesh*> m;
        std::vector<tm_energy*> e;
        for (size_t i = 0; i < meshes.size(); i++) {
            m.push_back(meshes[i]->h);
            e.push_back(estores[i]->h);
            nv.push_back(meshes[i]->nv());
        }
        check(tm_replicas_create(m.data(), e.data(), (int)m.size(), &h));
    }
    Replicas(const Replicas&) = delete;
    ~Replicas() { tm_replicas_destroy(h); }
    int size() const { return (int)nv.size(); }
    std::vector<double> per_replica(py::object v, const char* name) const
    {
        std::vector<double> out(size());
        if (py::isinstance<py::float_>(v) || py::isinstance<py::int_>(v)) {
            std::fill(out.begin(), out.end(), v.cast<double>());
            return out;
        }
        auto a = v.cast<std::vector<double>>();
        if ((int)a.size() != size()) throw std::runtime_error(std::string(name) + ": one value per replica");
        return a;
    }
};

std::atomic<uint64_t> g_seed_counter{0};
uint64_t fresh_seed()
{
    // like the reference's clock-seeded generator (flips.cpp:19-20)
    uint64_t t = (uint64_t)std::chrono::high_resolution_clock::now().time_since_epoch().count();
    return t ^ (0x9e3779b97f4a7c15ull * (++g_seed_counter));
}

}    // namespace

PYBIND11_MODULE(core, m)
{
    m.doc() = "trimem.core API on hand-written gfx950 HIP kernels (libtrimem_hip).";

    py::class_<HalfedgeHandle>(m, "HalfedgeHandle")
        .def(py::init())
        .def("idx", &HalfedgeHandle::idx)
        .def("is_valid", &HalfedgeHandle::is_valid);

    py::class_<TriMesh>(m, "TriMesh")
        .def(py::init([]() { return std::make_unique<TriMesh>(); }))
        .def(py::init(&from_points_cells), py::arg("points"), py::arg("faces"))
        .def("fv_indices", &fv_indices)
        .def("points", &points_view)
        .def("n_vertices", [](TriMesh& t) { return t.nv(); })
        .def("n_edges", [](TriMesh& t) { return t.ne(); })
        .def("n_faces", [](TriMesh& t) { return t.nf(); })
        .def("_handle", [](TriMesh& t) { return (uintptr_t)t.h; }, "the tm_mesh* of the C ABI")
        .def("halfedge_handle", [](TriMesh& t, int i) { return HalfedgeHandle(i >= 0 && i < 2 * t.ne() ? i : -1); },
             py::arg("edge"))
        .def("opposite_halfedge_handle",
             [](TriMesh&, const HalfedgeHandle& he) { return HalfedgeHandle(he.h >= 0 ? (he.h ^ 1) : -1); },
             py::arg("edge"))
        .def("next_halfedge_handle",
             [](TriMesh& t, const HalfedgeHandle& he) {
                 int o = -1;
                 check(tm_mesh_next_halfedge(t.h, he.h, &o));
                 return HalfedgeHandle(o);
             },
             py::arg("edge"));

    m.def("read_mesh", &read_mesh, py::arg("fname"));

    py::class_<VertexProperties>(m, "VertexProperties")
        .def(py::init())
        .def_readwrite("area", &VertexProperties::area)
        .def_readwrite("volume", &VertexProperties::volume)
        .def_readwrite("curvature", &VertexProperties::curvature)
        .def_readwrite("bending", &VertexProperties::bending)
        .def_readwrite("tethering", &VertexProperties::tethering)
        .def_readwrite("repulsion", &VertexProperties::repulsion);

    // mesh utilities on one halfedge (bindings.cpp:244-407), evaluated by the
    // same device functions the flip kernels use
    auto geom = [](TriMesh& t, const HalfedgeHandle& he) {
        std::vector<double> o(47);
        check(tm_halfedge_geometry(t.h, he.h, o.data()));
        return o;
    };
    auto rows = [](const std::vector<double>& o, int off, int r) {
        py::array_t<double> a({r, 3});
        std::memcpy(a.mutable_data(), o.data() + off, sizeof(double) * 3 * r);
        return a;
    };
    m.def("edge_length", [geom](TriMesh& t, HalfedgeHandle& he) { return geom(t, he)[0]; }, py::arg("mesh"),
          py::arg("halfedge_handle"));
    m.def("edge_length_grad", [geom, rows](TriMesh& t, HalfedgeHandle& he) { return rows(geom(t, he), 1, 2); },
          py::arg("mesh"), py::arg("halfedge_handle"));
    m.def("face_area", [geom](TriMesh& t, HalfedgeHandle& he) { return geom(t, he)[7]; }, py::arg("mesh"),
          py::arg("halfedge_handle"));
    m.def("face_area_grad", [geom, rows](TriMesh& t, HalfedgeHandle& he) { return rows(geom(t, he), 8, 3); },
          py::arg("mesh"), py::arg("halfedge_handle"));
    m.def("face_volume", [geom](TriMesh& t, HalfedgeHandle& he) { return geom(t, he)[17]; }, py::arg("mesh"),
          py::arg("halfedge_handle"));
    m.def("face_volume_grad", [geom, rows](TriMesh& t, HalfedgeHandle& he) { return rows(geom(t, he), 18, 3); },
          py::arg("mesh"), py::arg("halfedge_handle"));
    m.def("dihedral_angle", [geom](TriMesh& t, HalfedgeHandle& he) { return geom(t, he)[27]; }, py::arg("mesh"),
          py::arg("halfedge_handle"));
    m.def("dihedral_angle_grad",
          [geom, rows](TriMesh& t, HalfedgeHandle& he) { return rows(geom(t, he), 28, 4); }, py::arg("mesh"),
          py::arg("halfedge_handle"));

    py::enum_<BondType>(m, "BondType")
        .value("Edge", BondType::Edge)
        .value("Area", BondType::Area)
        .value("None", BondType::None)
        .export_values();

    py::class_<BondParams>(m, "BondParams")
        .def(py::init())
        .def_readwrite("lc0", &BondParams::lc0)
        .def_readwrite("lc1", &BondParams::lc1)
        .def_readwrite("a0", &BondParams::a0)
        .def_readwrite("r", &BondParams::r)
        .def_readwrite("type", &BondParams::type);

    py::class_<SurfaceRepulsionParams>(m, "SurfaceRepulsionParams")
        .def(py::init())
        .def_readwrite("lc1", &SurfaceRepulsionParams::lc1)
        .def_readwrite("r", &SurfaceRepulsionParams::r)
        .def_readwrite("n_search", &SurfaceRepulsionParams::n_search)
        .def_readwrite("rlist", &SurfaceRepulsionParams::rlist)
        .def_readwrite("exclusion_level", &SurfaceRepulsionParams::exclusion_level);

    py::class_<ContinuationParams>(m, "ContinuationParams")
        .def(py::init())
        .def_readwrite("delta", &ContinuationParams::delta)
        .def_readwrite("lam", &ContinuationParams::lam);

    py::class_<EnergyParams>(m, "EnergyParams")
        .def(py::init())
        .def_readwrite("kappa_b", &EnergyParams::kappa_b)
        .def_readwrite("kappa_a", &EnergyParams::kappa_a)
        .def_readwrite("kappa_v", &EnergyParams::kappa_v)
        .def_readwrite("kappa_c", &EnergyParams::kappa_c)
        .def_readwrite("kappa_t", &EnergyParams::kappa_t)
        .def_readwrite("kappa_r", &EnergyParams::kappa_r)
        .def_readwrite("area_frac", &EnergyParams::area_frac)
        .def_readwrite("volume_frac", &EnergyParams::volume_frac)
        .def_readwrite("curvature_frac", &EnergyParams::curvature_frac)
        .def_readwrite("bond_params", &EnergyParams::bond_params)
        .def_readwrite("repulse_params", &EnergyParams::repulse_params)
        .def_readwrite("continuation_params", &EnergyParams::continuation_params);

    py::class_<EnergyManager>(m, "EnergyManager")
        .def(py::init<TriMesh&, const EnergyParams&>(), py::arg("mesh"), py::arg("eparams"))
        .def("_handle", [](EnergyManager& e) { return (uintptr_t)e.h; }, "the tm_energy* of the C ABI")
        .def("properties", &EnergyManager::properties, py::arg("mesh"))
        .def("energy", &EnergyManager::energy, py::arg("mesh"))
        .def("energy", &EnergyManager::energy_props, py::arg("vprops"))
        .def("gradient", &EnergyManager::gradient, py::arg("mesh"))
        .def("update_reference_properties", [](EnergyManager& e) { check(tm_energy_update_reference_properties(e.h)); })
        .def("update_repulsion", [](EnergyManager& e, TriMesh& t) { check(tm_energy_update_repulsion(e.h, t.h)); })
        .def("print_info",
             [](EnergyManager& e, TriMesh& t) {
                 // energy.cpp:126-151
                 tm_props p;
                 check(tm_energy_properties(e.h, t.h, &p));
                 EnergyParams ep = e.eparams();
                 VertexProperties ip = e.initial_props();
                 double lam = ep.continuation_params.lam;
                 double ra = (1.0 - lam * (1.0 - ep.area_frac)) * ip.area;
                 double rv = (1.0 - lam * (1.0 - ep.volume_frac)) * ip.volume;
                 double rc = (1.0 - lam * (1.0 - ep.curvature_frac)) * ip.curvature;
                 auto sq = [](double x) { return x * x; };
                 double ea = ep.kappa_a * sq(p.area / ra - 1.0), ev = ep.kappa_v * sq(p.volume / rv - 1.0);
                 double ec = ep.kappa_c * sq(p.curvature / rc - 1.0);
                 double total;
                 check(tm_energy_energy_props(e.h, &p, &total));
                 std::ostringstream o;
                 o << "----- EnergyManager info\nreference properties:\n"
                   << "  area:      " << ra << "\n  volume:    " << rv << "\n  curvature: " << rc << "\n"
                   << "current properties:\n  area:      " << p.area << "\n  volume:    " << p.volume
                   << "\n  curvature: " << p.curvature << "\nenergies:\n  area:      " << ea
                   << "\n  volume:    " << ev << "\n  area diff: " << ec << "\n  bending:   "
                   << ep.kappa_b * p.bending << "\n  tether:    " << ep.kappa_t * p.tethering
                   << "\n  repulsion: " << ep.kappa_r * p.repulsion << "\n  total:     " << total << "\n";
                 py::print(o.str());
             },
             py::arg("mesh"))
        .def_property_readonly("eparams", &EnergyManager::eparams)
        .def_property_readonly("initial_props", &EnergyManager::initial_props)
        .def_property(
            "precision",
            [](EnergyManager& e) -> std::string {
                int p = 0;
                check(tm_energy_get_precision(e.h, &p));
                return p == TM_PREC_MIXED ? "mixed" : "fp64";
            },
            [](EnergyManager& e, const std::string& s) {
                if (s == "fp64") check(tm_energy_set_precision(e.h, TM_PREC_FP64));
                else if (s == "mixed") check(tm_energy_set_precision(e.h, TM_PREC_MIXED));
                else throw py::value_error("precision must be 'fp64' or 'mixed'");
            })
        .def_property(
            "tuning",
            [](EnergyManager& e) {
                tm_tuning t;
                check(tm_energy_get_tuning(e.h, &t));
                return tuning_dict(t);
            },
            [](EnergyManager& e, const py::dict& d) {
                tm_tuning t = tuning_of(d);
                check(tm_energy_set_tuning(e.h, &t));
            })
        .def(
            "autotune",
            [](EnergyManager& e, TriMesh& t) {
                tm_tuning out;
                double ms_rep = 0, ms_p2 = 0;
                check(tm_energy_autotune(e.h, t.h, &out, &ms_rep, &ms_p2));
                py::dict d = tuning_dict(out);
                d["rep_ms"] = ms_rep;
                d["pass2_ms"] = ms_p2;
                return d;
            },
            py::arg("mesh"));

    m.def(
        "flip",
        [](TriMesh& t, EnergyManager& e, double ratio, py::object seed) {
            uint64_t s = seed.is_none() ? fresh_seed() : seed.cast<uint64_t>();
            int acc = 0;
            check(tm_flip_serial(e.h, t.h, ratio, s, &acc));
            return acc;
        },
        py::arg("mesh"), py::arg("estore"), py::arg("flip_ratio"), py::kw_only(), py::arg("seed") = py::none());

    m.def(
        "pflip",
        [](TriMesh& t, EnergyManager& e, double ratio, py::object seed, int batch) {
            uint64_t s = seed.is_none() ? fresh_seed() : seed.cast<uint64_t>();
            int acc = 0;
            check(tm_flip_parallel(e.h, t.h, ratio, s, batch, &acc));
            return acc;
        },
        py::arg("mesh"), py::arg("estore"), py::arg("flip_ratio"), py::kw_only(), py::arg("seed") = py::none(),
        py::arg("batch") = 0);

    m.def(
        "flip_injected",
        [](TriMesh& t, EnergyManager& e, py::array_t<int, py::array::c_style | py::array::forcecast> proposals,
           py::array_t<double, py::array::c_style | py::array::forcecast> uniforms) {
            int acc = 0, used = 0;
            check(tm_flip_serial_injected(e.h, t.h, (int)proposals.size(), proposals.data(), uniforms.data(),
                                          (int)uniforms.size(), &acc, &used));
            return py::make_tuple(acc, used);
        },
        py::arg("mesh"), py::arg("estore"), py::arg("proposals"), py::arg("uniforms"));

    py::class_<Point>(m, "Point")
        .def_readonly("x", &Point::x)
        .def_readonly("y", &Point::y)
        .def_readonly("z", &Point::z)
        .def("__array__", [](const Point& p, py::args, py::kwargs) {
            py::array_t<double> a(3);
            a.mutable_at(0) = p.x; a.mutable_at(1) = p.y; a.mutable_at(2) = p.z;
            return a;
        });

    py::class_<NeighbourList>(m, "NeighbourList")
        .def("distance_matrix", &NeighbourList::distance_matrix, py::arg("mesh"), py::arg("rdist"))
        .def("point_distances", &NeighbourList::point_distances, py::arg("mesh"), py::arg("pid"), py::arg("rdist"));

    py::class_<DeviceHMC>(m, "DeviceHMC")
        .def(py::init([](TriMesh& t, EnergyManager* e, const std::string& model) {
                 return std::make_unique<DeviceHMC>(t, e, model);
             }),
             py::arg("mesh"), py::arg("estore") = nullptr, py::arg("model") = "energy", py::keep_alive<1, 2>(),
             py::keep_alive<1, 3>())
        .def("step", &DeviceHMC::step, py::kw_only(), py::arg("mass"), py::arg("time_step"), py::arg("num_integration_steps"),
            py::arg("temperature"), py::arg("momenta") = py::none(), py::arg("uniform") = -1.0,
            py::arg("seed") = 0, py::arg("step") = 0, py::arg("time_step_sq") = 0.0,
            py::arg("surrogate") = nullptr, py::arg("guarded") = false)
        .def("result", &DeviceHMC::result);

    py::class_<Surrogate>(m, "Surrogate")
        .def(py::init([](TriMesh& t, EnergyManager& e, uint64_t seed) { return std::make_unique<Surrogate>(t, e, seed); }),
             py::arg("mesh"), py::arg("estore"), py::arg("seed") = 0, py::keep_alive<1, 2>(), py::keep_alive<1, 3>())
        .def_property(
            "params",
            [](Surrogate& s) {
                int n = 0;
                check(tm_surrogate_params(s.h, nullptr, &n));
                py::array_t<float> out(n);
                check(tm_surrogate_params(s.h, out.mutable_data(), &n));
                return out;
            },
            [](Surrogate& s, py::array_t<float, py::array::c_style | py::array::forcecast> p) {
                int n = 0;
                check(tm_surrogate_params(s.h, nullptr, &n));
                if ((int)p.size() != n) throw std::runtime_error("wrong number of parameters");
                check(tm_surrogate_set_params(s.h, p.data()));
            })
        .def_property_readonly("norm",
                               [](Surrogate& s) {
                                   py::array_t<double> mu(8), sig(8);
                                   double lbar = 0, sf = 0;
                                   check(tm_surrogate_norm(s.h, mu.mutable_data(), sig.mutable_data(), &lbar, &sf));
                                   py::dict d;
                                   d["mu"] = mu;
                                   d["sig"] = sig;
                                   d["lbar"] = lbar;
                                   d["sigma_F"] = sf;
                                   return d;
                               })
        .def("forces",
             [](Surrogate& s) {
                 py::array_t<double> out({(size_t)s.nv, (size_t)3});
                 check(tm_surrogate_forces(s.h, out.mutable_data()));
                 return out;
             })
        .def("loss_grad",
             [](Surrogate& s) {
                 int n = 0;
                 check(tm_surrogate_params(s.h, nullptr, &n));
                 py::array_t<float> g(n);
                 double loss = 0;
                 check(tm_surrogate_loss_grad(s.h, &loss, g.mutable_data()));
                 return py::make_tuple(loss, g);
             })
        .def(
            "activations",
            [](Surrogate& s, int which) {
                const size_t w = which == 0 ? 16 : 64;
                py::array_t<float> out({(size_t)s.nv, w});
                check(tm_surrogate_activations(s.h, which, out.mutable_data()));
                return out;
            },
            py::arg("which"))
        .def(
            "set_policy",
            [](Surrogate& s, int window, double min_acceptance, int probe_every) {
                check(tm_surrogate_set_policy(s.h, window, min_acceptance, probe_every));
            },
            py::arg("window") = 10, py::arg("min_acceptance") = 0.5, py::arg("probe_every") = 10)
        .def("policy_state",
             [](Surrogate& s) {
                 int fb = 0;
                 long long ns = 0, ne = 0, nsw = 0;
                 double acc = 0;
                 check(tm_surrogate_policy_state(s.h, &fb, &ns, &ne, &nsw, &acc));
                 py::dict d;
                 d["falling_back"] = (bool)fb;
                 d["surrogate_steps"] = ns;
                 d["exact_steps"] = ne;
                 d["switches"] = nsw;
                 d["recent_acceptance"] = acc;
                 return d;
             })
        .def("add_snapshot", [](Surrogate& s) { check(tm_surrogate_add_snapshot(s.h)); })
        .def("num_snapshots",
             [](Surrogate& s) {
                 int n = 0;
                 check(tm_surrogate_num_snapshots(s.h, &n));
                 return n;
             })
        .def(
            "train",
            [](Surrogate& s, int epochs, float lr, uint64_t seed) {
                double loss = 0;
                check(tm_surrogate_train(s.h, epochs, lr, seed, &loss));
                return loss;
            },
            py::arg("epochs"), py::arg("lr") = 1e-3f, py::arg("seed") = 0);

    py::class_<Replicas>(m, "Replicas")
        .def(py::init([](const std::vector<TriMesh*>& meshes, const std::vector<EnergyManager*>& estores) {
                 return std::make_unique<Replicas>(meshes, estores);
             }),
             py::arg("meshes"), py::arg("estores"), py::keep_alive<1, 2>(), py::keep_alive<1, 3>())
        .def("__len__", &Replicas::size)
        .def(
            "eval", [](Replicas& r, bool props, bool grad) {
                check(tm_replicas_eval(r.h, (props ? TM_EVAL_PROPS : 0) | (grad ? TM_EVAL_GRAD : 0)));
            },
            py::arg("properties") = true, py::arg("gradient") = true)
        .def(
            "pflip",
            [](Replicas& r, double ratio, const std::vector<uint64_t>& seeds) {
                if ((int)seeds.size() != r.size()) throw std::runtime_error("one seed per replica");
                std::vector<int> acc(r.size(), 0);
                check(tm_replicas_pflip(r.h, ratio, seeds.data(), acc.data()));
                return acc;
            },
            "pflip on every replica at once (replica r draws from seeds[r]); accepted flips per replica",
            py::arg("flip_ratio"), py::arg("seeds"))
        .def("globals",
             [](Replicas& r) {
                 py::array_t<double> out({(size_t)r.size(), (size_t)7});
                 check(tm_replicas_globals(r.h, out.mutable_data()));
                 return out;
             })
        .def(
            "gradient",
            [](Replicas& r, int i) {
                if (i < 0 || i >= r.size()) throw py::index_error("replica index out of range");
                py::array_t<double> out({(size_t)r.nv[i], (size_t)3});
                check(tm_replicas_gradient(r.h, i, out.mutable_data()));
                return out;
            },
            py::arg("replica"))
        .def(
            "hmc_setup",
            [](Replicas& r, py::object mass, py::object time_step, py::object temperature,
               std::vector<uint64_t> seeds, int num_integration_steps) {
                auto ms = r.per_replica(mass, "mass"), dt = r.per_replica(time_step, "time_step"),
                     T = r.per_replica(temperature, "temperature");
                if ((int)seeds.size() != r.size()) throw std::runtime_error("seeds: one value per replica");
                check(tm_replicas_hmc_setup(r.h, ms.data(), dt.data(), T.data(), seeds.data(),
                                            num_integration_steps));
            },
            py::kw_only(), py::arg("mass"), py::arg("time_step"), py::arg("temperature"), py::arg("seeds"),
            py::arg("num_integration_steps"))
        .def(
            "set_ladder",
            [](Replicas& r, std::vector<double> ladder, uint64_t seed) {
                if ((int)ladder.size() != r.size()) throw std::runtime_error("ladder: one value per replica");
                check(tm_replicas_set_ladder(r.h, ladder.data(), seed));
            },
            py::arg("ladder"), py::arg("seed") = 0)
        .def(
            "hmc_run",
            [](Replicas& r, int n_steps, int swap_every, bool use_graph) {
                check(tm_replicas_hmc_run(r.h, n_steps, swap_every, use_graph ? 1 : 0));
            },
            py::arg("n_steps"), py::arg("swap_every") = 0, py::arg("use_graph") = false)
        .def("swap", [](Replicas& r) { check(tm_replicas_swap(r.h)); })
        .def("stats",
             [](Replicas& r) {
                 const size_t R = (size_t)r.size();
                 py::array_t<long long> acc(R), steps(R), sacc(R), stried(R);
                 py::array_t<int> slot(R);
                 py::array_t<double> props({R, (size_t)6});
                 check(tm_replicas_stats(r.h, acc.mutable_data(), steps.mutable_data(), slot.mutable_data(),
                                         props.mutable_data(), sacc.mutable_data(), stried.mutable_data()));
                 py::dict d;
                 d["accepted"] = acc;
                 d["steps"] = steps;
                 d["slot_of"] = slot;
                 d["props"] = props;
                 d["swaps_accepted"] = sacc;
                 d["swaps_tried"] = stried;
                 return d;
             })
        .def("update_repulsion", [](Replicas& r) { check(tm_replicas_update_repulsion(r.h)); })
        .def_property(
            "tuning",
            [](Replicas& r) {
                tm_tuning t{};
                check(tm_replicas_get_tuning(r.h, &t));
                return tuning_dict(t);
            },
            [](Replicas& r, const py::dict& d) {
                tm_tuning t = tuning_of(d);
                check(tm_replicas_set_tuning(r.h, &t));
            })
        .def("autotune", [](Replicas& r) {
            tm_tuning out;
            double ms_rep = 0, ms_p2 = 0;
            check(tm_replicas_autotune(r.h, &out, &ms_rep, &ms_p2));
            py::dict d = tuning_dict(out);
            d["rep_ms"] = ms_rep;
            d["pass2_ms"] = ms_p2;
            return d;
        });

    m.def(
        "make_nlist",
        [](TriMesh& t, const EnergyParams& e) {
            auto nl = std::make_unique<NeighbourList>();
            tm_params p = to_c(e);
            check(tm_nlist_create(t.h, &p, &nl->h));
            return nl;
        },
        py::arg("mesh"), py::arg("eparams"));

    m.def("area", [](TriMesh& t) { double a; check(tm_area(t.h, &a)); return a; }, py::arg("mesh"));
    m.def("edges_length", [](TriMesh& t) { double a; check(tm_edges_length(t.h, &a)); return a; }, py::arg("mesh"));
    m.def(
        "avg_tri_props",
        [](TriMesh& t) {
            double a, l;
            check(tm_area(t.h, &a));
            check(tm_edges_length(t.h, &l));
            return std::make_tuple(a / t.nf(), l / t.ne());
        },
        py::arg("mesh"));

    // finite-difference gradient (bindings.cpp:25-50)
    m.def(
        "gradient",
        [](TriMesh& t, EnergyManager& e, py::array_t<double> grad, double eps) {
            check(tm_mesh_sync_host(t.h));
            double e0 = e.energy(t);
            auto g = grad.mutable_unchecked<2>();
            double* pts = tm_mesh_points(t.h);
            for (int i = 0; i < t.nv(); i++) {
                for (int j = 0; j < 3; j++) {
                    pts[3 * i + j] += eps;
                    g(i, j) = (e.energy(t) - e0) / eps;
                    pts[3 * i + j] -= eps;
                }
            }
        },
        py::arg("mesh"), py::arg("estore"), py::arg("gradient"), py::arg("epsilon") = 1.0e-6);

    m.def("device_count", []() { int n = 0; check(tm_device_count(&n)); return n; });
}
