// tm_capi.hip -- host side of libtrimem_hip: mesh construction, device
// buffers and the C ABI declared in include/trimem_hip.h.
//
// Every compute entry point enqueues gfx950 kernels on one library stream;
// there is no CPU compute path.  Host work is limited to building the
// half-edge arrays from a face list (OpenMesh add_face numbering), scalar
// parameter arithmetic, and copying inputs and results.
#include "../../include/trimem_hip.h"
#include "tm_internal.h"
#include "tm_surrogate.h"

#include <cfloat>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <memory>
#include <numeric>
#include <random>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

using namespace tmh;

// ------------------------------------------------------------- errors -----
namespace {

thread_local std::string g_err;

struct TmError : std::runtime_error {
    int code;
    TmError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHECK(expr)                                                                             \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            throw TmError(TM_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));        \
    } while (0)

#define API_BEGIN try {
#define API_END                                          \
    }                                                    \
    catch (const TmError& ex) {                          \
        g_err = ex.what();                               \
        return ex.code;                                  \
    }                                                    \
    catch (const std::exception& ex) {                   \
        g_err = ex.what();                               \
        return TM_E_INVALID;                             \
    }                                                    \
    return TM_OK;

// ------------------------------------------------------------ runtime -----
// Scratch of tm_mesh_sum_noalloc: the arrival counter of its reduction (the
// first kNoAllocCounterBytes) and one partial sum per block of 256 faces or
// edges, in memory the library owns for the life of the process (one copy on
// every GPU it runs on), so that entry point allocates nothing.  Host code
// never passes this name to a kernel: it stands for a host-side shadow, and
// the device address comes from hipGetSymbolAddress (g_rt.scratch).
constexpr int kNoAllocMaxBlocks = 32768;                     // at most 32768 * 256 faces or edges
constexpr size_t kNoAllocCounterBytes = 256;
constexpr int kNoAllocScratchDoubles = kNoAllocCounterBytes / sizeof(double) + kNoAllocMaxBlocks;
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];
constexpr int kMaxDevices = 64;

struct Runtime {
    bool init = false;
    int device = 0;
    hipStream_t stream = nullptr;
    void* scratch[kMaxDevices] = {};    // device address of `scratch` on each device
    hipEvent_t pass1_marks[2] = {};     // tm_mark_pass1: taken by the next properties pass
};
Runtime g_rt;

void ensure_runtime()
{
    if (g_rt.init) return;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) throw TmError(TM_E_DEVICE, "no HIP device available");
    if (g_rt.device < 0 || g_rt.device >= n || g_rt.device >= kMaxDevices)
        throw TmError(TM_E_DEVICE, "device index out of range");
    HIPCHECK(hipSetDevice(g_rt.device));
    HIPCHECK(hipStreamCreateWithFlags(&g_rt.stream, hipStreamNonBlocking));
    // the scratch array's address on this device, with the device current
    void* sp = nullptr;
    HIPCHECK(hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)));
    if (!sp) throw TmError(TM_E_DEVICE, "no device address for the library scratch");
    g_rt.scratch[g_rt.device] = sp;
    g_rt.init = true;
}

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    void alloc(size_t count)
    {
        if (count <= n && p) return;
        release();
        size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
        HIPCHECK(hipMalloc((void**)&p, bytes));
        n = count;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DevBuf() { release(); }
};

void h2d(void* d, const void* h, size_t bytes)
{
    if (bytes) HIPCHECK(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, g_rt.stream));
}
void d2h(void* h, const void* d, size_t bytes)
{
    if (bytes) HIPCHECK(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, g_rt.stream));
}
void sync() { HIPCHECK(hipStreamSynchronize(g_rt.stream)); }
void check_launch() { HIPCHECK(hipGetLastError()); }

}    // namespace

// --------------------------------------------------------------- mesh -----
struct tm_mesh {
    int nv = 0, nh = 0, nf = 0;
    double* points = nullptr;    // host (N,3), pinned when a device exists
    bool pinned = false;
    // host mirror of the topology
    std::vector<int> he_to, he_next, he_prev, he_face, v_he, f_he;
    bool host_stale = false;     // device topology changed (flips)
    bool device_ahead = false;   // device positions moved by a device-resident step
    unsigned topo_version = 0;   // bumped by every topology change (flips)
    bool has_boundary = false;
    // device state
    bool dev = false;
    DevBuf<double> x;
    DevBuf<int> d_to, d_next, d_prev, d_face, d_vhe, d_fhe, ring, meta, flag;
    // vertices evaluated by the half-edge walker: boundary, isolated or with
    // fewer than three neighbours (flips preserve this set on valid meshes)
    DevBuf<int> generic;
    int n_generic = 0;
    // Internal vertex order (Morton order of the creation-time positions):
    // perm[i] = user id of internal vertex i, iperm[u] = internal id of user
    // vertex u.  Topology and device arrays use internal ids; every id or
    // per-vertex row that crosses the API is mapped back to user ids.
    std::vector<int> perm, iperm;
    DevBuf<int> d_perm, d_iperm;
    DevBuf<double> stage;    // user-order rows for uploads and downloads

    MeshView view() const
    {
        return MeshView{x.p, d_to.p, d_next.p, d_prev.p, d_face.p, d_vhe.p, d_fhe.p, ring.p, meta.p, nv, nh, nf};
    }
    FlipMeshView flip_view()
    {
        return FlipMeshView{x.p, d_to.p, d_next.p, d_prev.p, d_face.p, d_vhe.p, d_fhe.p, nv, nh, nf};
    }
    ~tm_mesh()
    {
        if (points) {
            if (pinned) (void)hipHostFree(points);
            else std::free(points);
        }
    }
};

namespace {

// Face list -> half-edge arrays with OpenMesh's numbering
// (PolyConnectivity::add_face as used by from_points_cells, mesh_py.cpp:53-67):
// face ids outside [0, n) are dropped, faces with fewer than three valid ids
// are skipped, a new edge e gets halfedge 2e = v_i -> v_{i+1}, the face's
// halfedge is v_2 -> v_0, and faces that would create a complex vertex (one
// whose fan is already closed) or a complex edge (a halfedge that already has
// a face) are rejected.  Remaining boundary halfedges are linked into loops
// and boundary vertices point at their boundary outgoing halfedge.
void build_topology(tm_mesh& M, const int* faces, int m)
{
    const int n = M.nv;
    M.v_he.assign(n, -1);
    M.he_to.clear(); M.he_next.clear(); M.he_prev.clear(); M.he_face.clear(); M.f_he.clear();
    if (n < 3 || m == 0) return;
    std::unordered_map<uint64_t, int> he_of;
    he_of.reserve((size_t)m * 4);
    auto key = [](int a, int b) { return ((uint64_t)(uint32_t)a << 32) | (uint32_t)b; };
    std::vector<int> n_out(n, 0), n_bout(n, 0);
    for (int f = 0; f < m; f++) {
        int vh[3], nvh = 0;
        for (int j = 0; j < 3; j++) {
            int c = faces[3 * f + j];
            if (c >= 0 && c < n) vh[nvh++] = c;
        }
        if (nvh < 3) continue;
        int hh[3];
        bool ok = true;
        for (int i = 0; i < 3 && ok; i++) {
            int a = vh[i], b = vh[(i + 1) % 3];
            if (n_out[a] > 0 && n_bout[a] == 0) ok = false;
            auto it = he_of.find(key(a, b));
            hh[i] = it == he_of.end() ? -1 : it->second;
            if (hh[i] >= 0 && M.he_face[hh[i]] >= 0) ok = false;
        }
        if (!ok) continue;
        int fid = (int)M.f_he.size();
        for (int i = 0; i < 3; i++) {
            if (hh[i] >= 0) continue;
            int a = vh[i], b = vh[(i + 1) % 3];
            int e = (int)M.he_to.size() / 2;
            M.he_to.push_back(b); M.he_to.push_back(a);
            M.he_next.push_back(-1); M.he_next.push_back(-1);
            M.he_prev.push_back(-1); M.he_prev.push_back(-1);
            M.he_face.push_back(-1); M.he_face.push_back(-1);
            he_of[key(a, b)] = 2 * e;
            he_of[key(b, a)] = 2 * e + 1;
            hh[i] = 2 * e;
            n_out[a]++; n_bout[a]++;
            n_out[b]++; n_bout[b]++;
            if (M.v_he[a] < 0) M.v_he[a] = 2 * e;
            if (M.v_he[b] < 0) M.v_he[b] = 2 * e + 1;
        }
        for (int i = 0; i < 3; i++) {
            M.he_face[hh[i]] = fid;
            M.he_next[hh[i]] = hh[(i + 1) % 3];
            M.he_prev[hh[(i + 1) % 3]] = hh[i];
            n_bout[vh[i]]--;
        }
        M.f_he.push_back(hh[2]);
    }
    const int nh = (int)M.he_to.size();
    std::vector<int> bout(n, -1);
    for (int h = 0; h < nh; h++)
        if (M.he_face[h] < 0) bout[M.he_to[h ^ 1]] = h;
    for (int h = 0; h < nh; h++) {
        if (M.he_face[h] < 0) {
            int nx = bout[M.he_to[h]];
            M.he_next[h] = nx;
            if (nx >= 0) M.he_prev[nx] = h;
        }
    }
    M.has_boundary = false;
    for (int v = 0; v < n; v++) {
        if (bout[v] >= 0) {
            M.v_he[v] = bout[v];
            M.has_boundary = true;
        }
    }
    M.nh = nh;
    M.nf = (int)M.f_he.size();
}

// Morton (Z-order) key of each vertex on a 1024^3 grid over the bounding box,
// ties broken by id; vertices close in space get close internal ids, which
// keeps the gathers of one wavefront in few cache lines.
void reorder_vertices(tm_mesh& M)
{
    const int n = M.nv;
    M.perm.resize(n);
    M.iperm.resize(n);
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int v = 0; v < n; v++)
        for (int k = 0; k < 3; k++) {
            double c = M.points[3 * v + k];
            if (c < lo[k]) lo[k] = c;
            if (c > hi[k]) hi[k] = c;
        }
    auto spread = [](uint64_t x) {
        x &= 0x3ff;
        x = (x | (x << 16)) & 0x030000ffull;
        x = (x | (x << 8)) & 0x0300f00full;
        x = (x | (x << 4)) & 0x030c30c3ull;
        x = (x | (x << 2)) & 0x09249249ull;
        return x;
    };
    std::vector<uint64_t> key(n);
    for (int v = 0; v < n; v++) {
        uint64_t code = 0;
        for (int k = 0; k < 3; k++) {
            double span = hi[k] - lo[k];
            double t = span > 0 ? (M.points[3 * v + k] - lo[k]) / span : 0.0;
            if (!(t >= 0)) t = 0;    // also catches NaN
            if (t > 1) t = 1;
            code |= spread((uint64_t)(t * 1023.0)) << k;
        }
        key[v] = (code << 32) | (uint32_t)v;
    }
    std::sort(key.begin(), key.end());
    for (int i = 0; i < n; i++) {
        M.perm[i] = (int)(key[i] & 0xffffffffu);
        M.iperm[M.perm[i]] = i;
    }
    for (int h = 0; h < M.nh; h++) M.he_to[h] = M.iperm[M.he_to[h]];
    std::vector<int> vhe(n);
    for (int i = 0; i < n; i++) vhe[i] = M.v_he[M.perm[i]];
    M.v_he.swap(vhe);
}

void ensure_device(tm_mesh& M)
{
    ensure_runtime();
    if (M.dev) return;
    M.x.alloc((size_t)M.nv * 3);
    M.d_to.alloc(M.nh);
    M.d_next.alloc(M.nh);
    M.d_prev.alloc(M.nh);
    M.d_face.alloc(M.nh);
    M.d_vhe.alloc(M.nv);
    M.d_fhe.alloc(M.nf);
    M.ring.alloc((size_t)kMaxValence * M.nv);
    M.meta.alloc(M.nv);
    M.d_perm.alloc(M.nv);
    M.d_iperm.alloc(M.nv);
    M.stage.alloc((size_t)M.nv * 6);
    h2d(M.d_perm.p, M.perm.data(), sizeof(int) * M.nv);
    h2d(M.d_iperm.p, M.iperm.data(), sizeof(int) * M.nv);
    M.flag.alloc(4);
    h2d(M.d_to.p, M.he_to.data(), sizeof(int) * M.nh);
    h2d(M.d_next.p, M.he_next.data(), sizeof(int) * M.nh);
    h2d(M.d_prev.p, M.he_prev.data(), sizeof(int) * M.nh);
    h2d(M.d_face.p, M.he_face.data(), sizeof(int) * M.nh);
    h2d(M.d_vhe.p, M.v_he.data(), sizeof(int) * M.nv);
    h2d(M.d_fhe.p, M.f_he.data(), sizeof(int) * M.nf);
    std::vector<int> gl;
    for (int v = 0; v < M.nv; v++) {
        int h0 = M.v_he[v];
        if (h0 < 0 || M.he_face[h0] < 0) {
            gl.push_back(v);
            continue;
        }
        int n = 0, h = h0;
        do {
            n++;
            h = M.he_next[h ^ 1];
        } while (h != h0 && n < 4);
        if (n < 3) gl.push_back(v);
    }
    M.n_generic = (int)gl.size();
    M.generic.alloc(std::max<size_t>(gl.size(), 1));
    h2d(M.generic.p, gl.data(), sizeof(int) * gl.size());
    HIPCHECK(hipMemsetAsync(M.flag.p, 0, sizeof(int) * 4, g_rt.stream));
    launch_build_ring(M.view(), M.ring.p, M.meta.p, M.flag.p, g_rt.stream);
    check_launch();
    int ovf = 0;
    d2h(&ovf, M.flag.p, sizeof(int));
    sync();
    if (ovf) throw TmError(TM_E_UNSUPPORTED, "vertex valence exceeds the ring-table width (32)");
    M.dev = true;
}

void push_points(tm_mesh& M)
{
    ensure_device(M);
    M.device_ahead = false;
    if (M.nv == 0) return;
    h2d(M.stage.p, M.points, sizeof(double) * 3 * M.nv);
    launch_gather_rows(M.stage.p, M.d_perm.p, M.x.p, M.nv, 3, g_rt.stream);
    check_launch();
}

// Host-facing calls take positions from the host array (the reference's
// points() view, mesh_py.cpp:96-105).  When a device-resident step has moved
// the device copy since, the device holds the newer state and is used as it
// is; the host array is refreshed only when it is read (tm_mesh_sync_host).
void upload_points(tm_mesh& M)
{
    ensure_device(M);
    if (!M.device_ahead) push_points(M);
}

// rows of an internal-order device array -> user order in the stage buffer
// -> host
void download_rows(tm_mesh& M, const double* d_src, int width, double* host)
{
    if (M.nv == 0) return;
    launch_gather_rows(d_src, M.d_iperm.p, M.stage.p, M.nv, width, g_rt.stream);
    check_launch();
    d2h(host, M.stage.p, sizeof(double) * width * M.nv);
}

void pull_points(tm_mesh& M)
{
    download_rows(M, M.x.p, 3, M.points);
    sync();
    M.device_ahead = false;
}

void refresh_host_topology(tm_mesh& M)
{
    if (!M.host_stale) return;
    d2h(M.he_to.data(), M.d_to.p, sizeof(int) * M.nh);
    d2h(M.he_next.data(), M.d_next.p, sizeof(int) * M.nh);
    d2h(M.he_prev.data(), M.d_prev.p, sizeof(int) * M.nh);
    d2h(M.he_face.data(), M.d_face.p, sizeof(int) * M.nh);
    d2h(M.v_he.data(), M.d_vhe.p, sizeof(int) * M.nv);
    d2h(M.f_he.data(), M.d_fhe.p, sizeof(int) * M.nf);
    sync();
    M.host_stale = false;
}

}    // namespace

// ---------------------------------------------------------- nlist ---------
struct tm_nlist {
    int kind = 0;    // 0 cell, 1 verlet
    int exclusion = 0;
    double rlist = 0.1;
    NListView v{};
    DevBuf<int> cell_start, cell_verts, vert_cell, counts, fill, scan_tmp, adj_start, adj, adj_counts;
    DevBuf<int> cand_count, cand_start, cand;
    DevBuf<int2> vcand;
    DevBuf<unsigned long long> excl_mask;
    unsigned mask_topo = ~0u;    // topology version the masks were built for
    // pair cache of the query (PairCache): recorded lists, valid for this
    // list and topology once a recording launch has been queued
    DevBuf<double> pc_x, pc_disp;
    DevBuf<int> pc_cell, pc_ent;
    DevBuf<unsigned> pc_cnt;
    bool pc_ok = false;
    // distance-matrix scratch
    DevBuf<int> dm_counts, dm_off;
    DevBuf<double> dm_d;
    DevBuf<int> dm_i, dm_j;
    DevBuf<double> box_part, pd_dist;
    DevBuf<int> pd_j, pd_count;
    int nv = 0;
    NListView view() const { return v; }
};

namespace {

// (Re)build the exclusion masks of a cell list when the mesh topology has
// changed since they were built; the view carries them only while valid.
void ensure_masks(tm_nlist& L, tm_mesh& M)
{
    if (L.kind != 0 || !L.excl_mask.p || M.nv == 0) return;
    if (L.mask_topo != M.topo_version) {
        NListView nv = L.v;
        nv.excl_mask = nullptr;
        launch_excl_mask(M.view(), nv, L.excl_mask.p, g_rt.stream);
        check_launch();
        L.mask_topo = M.topo_version;
        L.pc_ok = false;    // exclusion may have changed
    }
    L.v.excl_mask = L.excl_mask.p;
}

void scan(const int* in, int* out, int n, DevBuf<int>& tmp)
{
    tmp.alloc(scan_block_sums_size(n) + 1);
    launch_exclusive_scan(in, out, n, tmp.p, g_rt.stream);
    check_launch();
}

// CellList ctor (cell_list.h:43-87) and, for kind 1, VerletList ctor
// (verlet_list.h:18-35), on the mesh's current device positions.
void build_nlist(tm_nlist& L, tm_mesh& M)
{
    if (!(L.rlist > 0) || !std::isfinite(L.rlist)) throw TmError(TM_E_INVALID, "rlist must be positive");
    const int n = M.nv;
    L.nv = n;
    const int nbox = box_scratch_doubles();
    L.box_part.alloc(nbox);
    launch_box(M.x.p, n, L.box_part.p, g_rt.stream);
    check_launch();
    std::vector<double> part(nbox);
    d2h(part.data(), L.box_part.p, sizeof(double) * nbox);
    sync();
    double tight[6] = {DBL_MAX, DBL_MAX, DBL_MAX, -DBL_MAX, -DBL_MAX, -DBL_MAX};    // the vertices' own box
    for (int b = 0; b < nbox / 6; b++)
        for (int k = 0; k < 3; k++) {
            tight[k] = std::fmin(tight[k], part[6 * b + k]);
            tight[3 + k] = std::fmax(tight[3 + k], part[6 * b + 3 + k]);
        }
    // the reference's box (cell_list.h:46-56): its bounds start at min =
    // (+max, 0, 0) and max = (-1, 0, 0), so it spans 0 in y and z
    double box[6];
    for (int k = 0; k < 3; k++) {
        box[k] = k == 0 ? tight[0] : std::min(0.0, tight[k]);
        box[3 + k] = std::max(k == 0 ? -1.0 : 0.0, tight[3 + k]);
    }
    const double box_eps = 1.0e-6;
    NListView v{};
    v.kind = 0;
    v.exclusion = L.exclusion;
    long long ncells = 1;
    for (int k = 0; k < 3; k++) {
        if (n == 0) {
            v.box_min[k] = 0.0;
            v.r_list[k] = 1.0;
            v.lo[k] = 0;
            v.shape[k] = 1;
            continue;
        }
        const double dim = box[3 + k] - box[k] + box_eps;
        const double q = dim / L.rlist;
        if (!(q < 1073741824.0))
            throw TmError(TM_E_UNSUPPORTED, "cell grid too fine along axis " + std::to_string(k) + " (" +
                                                std::to_string(q) + " cells); coordinates must be finite");
        const int nk = int(q);
        const int gk = (nk > 0) ? nk : 1;
        v.box_min[k] = box[k];
        v.r_list[k] = (nk > 0) ? dim / nk : dim;
        // global cells of the lowest and highest vertex; the cell index is
        // monotone in the coordinate, so they bound every vertex's cell
        auto cell = [&](double p) { return std::min(std::max(int((p - box[k]) / v.r_list[k]), 0), gk - 1); };
        v.lo[k] = cell(tight[k]);
        v.shape[k] = cell(tight[3 + k]) - v.lo[k] + 1;
        ncells *= v.shape[k];
    }
    if (ncells > (1ll << 28))
        throw TmError(TM_E_UNSUPPORTED, "occupied cell sub-grid too large (" + std::to_string(ncells) +
                                            " cells): the mesh spans more than 2^28 cells of size rlist");
    v.strides[0] = 1;
    v.strides[1] = v.shape[0];
    v.strides[2] = v.shape[0] * v.shape[1];
    const int nc = (int)ncells;
    L.counts.alloc(nc + 1);
    L.cell_start.alloc(nc + 1);
    L.fill.alloc(nc);
    L.vert_cell.alloc(n);
    L.cell_verts.alloc(n);
    const bool cells = L.kind == 0 && n > 0;    // the per-cell candidate lists
    if (cells) L.cand_count.alloc(n + 1);
    launch_fill_ints(L.counts.p, nc + 1, 0, L.fill.p, nc, 0, cells ? L.cand_count.p + n : nullptr, cells ? 1 : 0, 0,
                     g_rt.stream);
    check_launch();
    launch_cell_ids(M.x.p, n, v, L.vert_cell.p, L.counts.p, g_rt.stream);
    check_launch();
    scan(L.counts.p, L.cell_start.p, nc + 1, L.scan_tmp);
    launch_cell_scatter(L.vert_cell.p, n, L.cell_start.p, L.fill.p, L.cell_verts.p, g_rt.stream);
    check_launch();
    launch_cell_sort_segments(L.vert_cell.p, n, L.cell_start.p, M.d_perm.p, L.cell_verts.p, g_rt.stream);
    check_launch();
    v.cell_start = L.cell_start.p;
    v.cell_verts = L.cell_verts.p;
    v.vert_cell = L.vert_cell.p;
    v.perm = M.d_perm.p;
    if (L.kind == 0 && n > 0) {
        if ((long long)27 * n > INT_MAX) throw TmError(TM_E_UNSUPPORTED, "mesh too large for the candidate table");
        L.cand_start.alloc(n + 1);
        launch_cand_count(v, n, L.cand_count.p, g_rt.stream);
        check_launch();
        scan(L.cand_count.p, L.cand_start.p, n + 1, L.scan_tmp);
        // a vertex is in the stencil lists of at most 27 cells: 27 N bounds
        // the total without reading it back
        L.cand.alloc((size_t)27 * n);
        launch_cand_fill(v, n, L.cand_start.p, L.cand.p, g_rt.stream);
        check_launch();
        L.vcand.alloc(n);
        launch_cand_vrange(v, n, L.cand_start.p, L.vcand.p, g_rt.stream);
        check_launch();
        v.vcand = L.vcand.p;
        v.cand = L.cand.p;
        v.cand_start = L.cand_start.p;
    }
    if (L.kind == 1) {
        L.adj_counts.alloc(n + 1);
        L.adj_start.alloc(n + 1);
        HIPCHECK(hipMemsetAsync(L.adj_counts.p, 0, sizeof(int) * (n + 1), g_rt.stream));
        launch_verlet_count(M.view(), v, L.rlist, L.adj_counts.p, g_rt.stream);
        check_launch();
        scan(L.adj_counts.p, L.adj_start.p, n + 1, L.scan_tmp);
        int total = 0;
        d2h(&total, L.adj_start.p + n, sizeof(int));
        sync();
        L.adj.alloc(total);
        launch_verlet_fill(M.view(), v, L.rlist, L.adj_start.p, L.adj.p, g_rt.stream);
        check_launch();
        v.kind = 1;
        v.adj_start = L.adj_start.p;
        v.adj = L.adj.p;
    }
    L.v = v;
    if (L.kind == 0 && n > 0) {
        L.excl_mask.alloc((size_t)2 * n);
        L.mask_topo = ~0u;
        ensure_masks(L, M);
        L.pc_x.alloc((size_t)3 * n);
        L.pc_cell.alloc(n);
        L.pc_cnt.alloc(n);
        L.pc_ent.alloc((size_t)kPairLanes * kPairCap * n);
        const int nd = pass1_blocks(n);    // one displacement partial per pass-1 block
        L.pc_disp.alloc(nd);
        // cell -1: no replay; displacement bytes 0x7F (~1e306): a recording first
        launch_fill_ints(L.pc_cell.p, n, -1, reinterpret_cast<int*>(L.pc_disp.p), 2 * nd, 0x7F7F7F7F,
                         nullptr, 0, 0, g_rt.stream);
        check_launch();
        L.pc_ok = false;
    }
    // no wait here: everything that reads the list is queued behind it on the
    // library stream (the host already waited once, for the bounding box)
}

void check_params(const tm_params& p)
{
    if (p.bond_type != 0 && p.bond_type != 1) throw TmError(TM_E_INVALID, "Unknown bond potential");
    if (p.rep_nlist != 0 && p.rep_nlist != 1) throw TmError(TM_E_INVALID, "Unknown neighbour search algorithm.");
    if (p.rep_exclusion < 0 || p.rep_exclusion > 2) throw TmError(TM_E_INVALID, "Unsupported exclusion level");
}

tm_nlist* make_nlist(tm_mesh& M, const tm_params& p)
{
    check_params(p);
    auto L = std::make_unique<tm_nlist>();
    L->kind = p.rep_nlist;
    L->exclusion = p.rep_exclusion;
    L->rlist = p.rep_rlist;
    build_nlist(*L, M);
    return L.release();
}

}    // namespace

// --------------------------------------------------------- energy ---------
struct tm_energy {
    tm_params params{};
    tm_props initial{};
    int precision = TM_PREC_FP64;
    tm_nlist* nlist = nullptr;
    int nv = 0;
    DevBuf<double> q, partials, grad, vprops, gprops;
    DevBuf<int> pair_cnt;
    DevBuf<double> rep_v, rep_g;
    DevBuf<double> theta;    // pass 1's dihedral angles for pass 2 (kRingPad per vertex, float or double)
    DevBuf<Globals> globals;
    DevBuf<int> flip_result, proposals, pf_status;
    DevBuf<double> uniforms, pf_dprops, pf_state;
    double pf_last[8] = {};    // pf_state at the end of the last parallel sweep (tm_energy_pflip_state)
    DevBuf<unsigned long long> pf_owner;
    DevBuf<SerialState> sf_state;
    DevBuf<int> sf_int;
    DevBuf<double> sf_dbl;
    DevBuf<unsigned long long> sf_vstamp, sf_estamp, sf_fstamp;
    Tuning tune;
    tm_nlist* spare = nullptr;    // the previous list, rebuilt in place by update_repulsion
    // bumped whenever something a replica group copies into its device
    // table changes: the neighbour list, the reference values, the precision
    unsigned gen = 0;
    ~tm_energy()
    {
        delete nlist;
        delete spare;
    }
};

namespace {

EParams eparams_of(const tm_energy& E)
{
    const tm_params& P = E.params;
    EParams p{};
    p.kappa_b = P.kappa_b; p.kappa_a = P.kappa_a; p.kappa_v = P.kappa_v;
    p.kappa_c = P.kappa_c; p.kappa_t = P.kappa_t; p.kappa_r = P.kappa_r;
    p.bond_lc0 = P.bond_lc0; p.bond_lc1 = P.bond_lc1; p.bond_a0 = P.bond_a0;
    p.bond_r = P.bond_r; p.bond_type = P.bond_type;
    p.rep_lc1 = P.rep_lc1; p.rep_r = P.rep_r; p.rep_exclusion = P.rep_exclusion;
    // interpolate_reference_properties (energy.cpp:30-45)
    double lam = P.cont_lambda;
    p.ref_area = (1.0 - lam * (1.0 - P.area_frac)) * E.initial.area;
    p.ref_volume = (1.0 - lam * (1.0 - P.volume_frac)) * E.initial.volume;
    p.ref_curvature = (1.0 - lam * (1.0 - P.curvature_frac)) * E.initial.curvature;
    return p;
}

void ensure_energy_buffers(tm_energy& E, const tm_mesh& M)
{
    if (E.nv == M.nv && E.q.p) return;
    E.nv = M.nv;
    E.q.alloc(M.nv);
    E.gprops.alloc((size_t)M.nv * 6);
    E.pair_cnt.alloc(M.nv);
    HIPCHECK(hipMemsetAsync(E.pair_cnt.p, 0, sizeof(int) * M.nv, g_rt.stream));
    E.rep_v.alloc(M.nv);
    E.rep_g.alloc((size_t)M.nv * 3);
    E.partials.alloc((size_t)pass1_blocks(M.nv) * 6);
    E.grad.alloc((size_t)M.nv * 3);
    E.theta.alloc((size_t)kRingPad * M.nv);
    E.globals.alloc(1);
    E.flip_result.alloc(4);
}

// update_repulsion (energy.cpp:57-61) without new device allocations: the
// list is built into the energy manager's spare, whose buffers remain from
// the rebuild before, and swapped in only when complete, so a failed build
// leaves the current list in place.
void rebuild_nlist(tm_energy& E, tm_mesh& M)
{
    check_params(E.params);
    if (!E.spare) E.spare = new tm_nlist();
    tm_nlist& L = *E.spare;
    L.kind = E.params.rep_nlist;
    L.exclusion = E.params.rep_exclusion;
    L.rlist = E.params.rep_rlist;
    build_nlist(L, M);
    std::swap(E.nlist, E.spare);
    E.gen++;
}

// The pair-cache buffers of the energy manager's cell list (x_snap null when
// it has none: a Verlet list, or no repulsion); the caller adds the
// displacement partials and the recording flag.
PairCache pair_cache_of(tm_energy& E, tm_mesh& M)
{
    PairCache pc{};
    tm_nlist& L = *E.nlist;
    if (L.kind == 0 && L.pc_x.p && L.nv == M.nv && E.params.rep_lc1 > 0) {
        const double skin = kPairSkin * E.params.rep_lc1;
        pc.x_snap = L.pc_x.p;
        pc.cell = L.pc_cell.p;
        pc.cnt = L.pc_cnt.p;
        pc.ent = L.pc_ent.p;
        pc.skin = skin;
        pc.limit2 = (kPairLimit * skin) * (kPairLimit * skin);
    }
    return pc;
}

PassBuffers buffers(tm_energy& E, tm_mesh& M)
{
    PassBuffers b{};
    b.q = E.q.p;
    b.gprops = E.gprops.p;
    b.pair_cnt = E.pair_cnt.p;
    b.rep_v = E.rep_v.p;
    b.rep_g = E.rep_g.p;
    b.partials = E.partials.p;
    b.globals = E.globals.p;
    b.theta = E.theta.p;
    b.list = M.generic.p;
    b.nlist = M.n_generic;
    tm_nlist& L = *E.nlist;
    if (E.tune.pair_cache) {
        b.pc = pair_cache_of(E, M);
        if (b.pc.x_snap) {
            b.pc.disp_part = L.pc_disp.p;
            b.pc.disp_n = pass1_blocks(M.nv);
            b.pc.force = L.pc_ok ? 0 : 1;
        }
    }
    return b;
}

void run_pass1(tm_energy& E, tm_mesh& M, int mixed, double* vprops)
{
    ensure_energy_buffers(E, M);
    if (E.nlist->nv != M.nv) throw TmError(TM_E_INVALID, "mesh does not match the energy manager");
    if (M.nv == 0) {
        HIPCHECK(hipMemsetAsync(E.globals.p, 0, sizeof(Globals), g_rt.stream));
        return;
    }
    ensure_masks(*E.nlist, M);
    const PassBuffers b = buffers(E, M);
    hipEvent_t marks[2] = {g_rt.pass1_marks[0], g_rt.pass1_marks[1]};
    g_rt.pass1_marks[0] = g_rt.pass1_marks[1] = nullptr;
    launch_pass1(M.view(), E.nlist->view(), eparams_of(E), mixed, b, vprops, g_rt.stream, E.tune, marks);
    check_launch();
    if (b.pc.x_snap) E.nlist->pc_ok = true;    // recorded by this launch, or still valid
}

void run_pass2(tm_energy& E, tm_mesh& M, int mixed)
{
    if (M.nv == 0) return;
    ensure_masks(*E.nlist, M);
    launch_pass2(M.view(), E.nlist->view(), eparams_of(E), mixed, buffers(E, M), E.grad.p, g_rt.stream, E.tune);
    check_launch();
}

Globals read_globals(tm_energy& E)
{
    Globals g;
    d2h(&g, E.globals.p, sizeof(Globals));
    sync();
    return g;
}

tm_props props_of(const Globals& g)
{
    return tm_props{g.props[0], g.props[1], g.props[2], g.props[3], g.props[4], g.props[5]};
}

// trimem_energy (kernel.h:110-122) on host-supplied properties
double energy_from(const tm_energy& E, const tm_props& pr)
{
    EParams p = eparams_of(E);
    double e = 0.0, d;
    d = pr.area / p.ref_area - 1.0;
    e += p.kappa_a * d * d;
    d = pr.volume / p.ref_volume - 1.0;
    e += p.kappa_v * d * d;
    d = pr.curvature / p.ref_curvature - 1.0;
    e += p.kappa_c * d * d;
    e += p.kappa_t * pr.tethering;
    e += p.kappa_r * pr.repulsion;
    e += p.kappa_b * pr.bending;
    return e;
}

void require(bool c, const char* msg)
{
    if (!c) throw TmError(TM_E_INVALID, msg);
}

}    // namespace

// ================================================================ C ABI ====
extern "C" {

const char* tm_last_error(void) { return g_err.c_str(); }

int tm_device_count(int* count)
{
    API_BEGIN
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
    API_END
}

int tm_set_device(int device)
{
    API_BEGIN
    require(!g_rt.init || g_rt.device == device, "device already selected");
    g_rt.device = device;
    API_END
}

int tm_mesh_create(const double* points, int n_points, const int* faces, int n_faces, tm_mesh** out)
{
    API_BEGIN
    require(out != nullptr, "null output");
    require(n_points >= 0 && n_faces >= 0, "negative size");
    auto M = std::make_unique<tm_mesh>();
    M->nv = n_points;
    size_t bytes = sizeof(double) * 3 * std::max(n_points, 1);
    // pinned host memory when a device is present (fast uploads); plain
    // memory otherwise so that meshes can be built on a host without a GPU
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) == hipSuccess && ndev > 0 &&
        hipHostMalloc((void**)&M->points, bytes, hipHostMallocDefault) == hipSuccess) {
        M->pinned = true;
    } else {
        (void)hipGetLastError();
        M->points = (double*)std::malloc(bytes);
        if (!M->points) throw TmError(TM_E_INVALID, "out of host memory");
    }
    if (n_points) std::memcpy(M->points, points, sizeof(double) * 3 * n_points);
    build_topology(*M, faces, n_faces);
    reorder_vertices(*M);
    *out = M.release();
    API_END
}

int tm_mesh_destroy(tm_mesh* mesh)
{
    delete mesh;
    return TM_OK;
}

int tm_mesh_counts(const tm_mesh* mesh, int* nv, int* ne, int* nf)
{
    API_BEGIN
    if (nv) *nv = mesh->nv;
    if (ne) *ne = mesh->nh / 2;
    if (nf) *nf = mesh->nf;
    API_END
}

double* tm_mesh_points(tm_mesh* mesh) { return mesh->points; }

int tm_mesh_fv_indices(tm_mesh* mesh, int* out)
{
    API_BEGIN
    refresh_host_topology(*mesh);
    for (int f = 0; f < mesh->nf; f++) {
        int h = mesh->f_he[f];
        out[3 * f] = mesh->perm[mesh->he_to[h]];
        h = mesh->he_next[h];
        out[3 * f + 1] = mesh->perm[mesh->he_to[h]];
        h = mesh->he_next[h];
        out[3 * f + 2] = mesh->perm[mesh->he_to[h]];
    }
    API_END
}

int tm_mesh_next_halfedge(tm_mesh* mesh, int he, int* out)
{
    API_BEGIN
    refresh_host_topology(*mesh);
    *out = (he >= 0 && he < mesh->nh) ? mesh->he_next[he] : -1;
    API_END
}

int tm_mesh_topology(tm_mesh* mesh, int* he_to, int* he_next, int* he_face, int* v_he, int* f_he)
{
    API_BEGIN
    refresh_host_topology(*mesh);
    if (he_to)
        for (int h = 0; h < mesh->nh; h++) he_to[h] = mesh->perm[mesh->he_to[h]];
    if (he_next) std::memcpy(he_next, mesh->he_next.data(), sizeof(int) * mesh->nh);
    if (he_face) std::memcpy(he_face, mesh->he_face.data(), sizeof(int) * mesh->nh);
    if (v_he)
        for (int u = 0; u < mesh->nv; u++) v_he[u] = mesh->v_he[mesh->iperm[u]];
    if (f_he) std::memcpy(f_he, mesh->f_he.data(), sizeof(int) * mesh->nf);
    API_END
}

double* tm_mesh_device_points(tm_mesh* mesh)
{
    try {
        ensure_device(*mesh);
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return nullptr;
    }
    return mesh->x.p;
}

int tm_mesh_upload_points(tm_mesh* mesh)
{
    API_BEGIN
    push_points(*mesh);
    sync();
    API_END
}

int tm_mesh_download_points(tm_mesh* mesh)
{
    API_BEGIN
    ensure_device(*mesh);
    pull_points(*mesh);
    API_END
}

int tm_mesh_sync_host(tm_mesh* mesh)
{
    API_BEGIN
    if (mesh->device_ahead) pull_points(*mesh);
    API_END
}

int tm_halfedge_geometry(tm_mesh* mesh, int he, double* out47)
{
    API_BEGIN
    require(he >= 0 && he < mesh->nh, "halfedge out of range");
    upload_points(*mesh);
    DevBuf<double> d;
    d.alloc(47);
    launch_halfedge_geometry(mesh->view(), he, d.p, g_rt.stream);
    check_launch();
    d2h(out47, d.p, sizeof(double) * 47);
    sync();
    API_END
}

static int sum_geometry(tm_mesh* mesh, int what, double* out)
{
    API_BEGIN
    upload_points(*mesh);
    int n = what == 0 ? mesh->nf : mesh->nh / 2;
    DevBuf<double> part, res;
    DevBuf<unsigned> cnt;
    part.alloc((n + 255) / 256 + 1);
    res.alloc(1);
    cnt.alloc(1);
    HIPCHECK(hipMemsetAsync(cnt.p, 0, sizeof(unsigned), g_rt.stream));
    HIPCHECK(hipMemsetAsync(res.p, 0, sizeof(double), g_rt.stream));
    if (n > 0) {
        launch_sum_geometry(mesh->view(), what, part.p, cnt.p, res.p, g_rt.stream);
        check_launch();
    }
    d2h(out, res.p, sizeof(double));
    sync();
    API_END
}

int tm_area(tm_mesh* mesh, double* out) { return sum_geometry(mesh, 0, out); }
int tm_edges_length(tm_mesh* mesh, double* out) { return sum_geometry(mesh, 1, out); }

// tm_area / tm_edges_length without allocating: the device-resident
// positions, the partials and counter in the library's scratch, the sum
// written to the caller's device double, nothing synchronised.
int tm_mesh_sum_noalloc(tm_mesh* mesh, int what, double* out)
{
    API_BEGIN
    require(mesh != nullptr && out != nullptr, "null argument");
    require(what == 0 || what == 1, "what must be 0 (face areas) or 1 (edge lengths)");
    ensure_runtime();
    require(mesh->dev, "the mesh is not on the device yet (upload its points first)");
    const int n = what == 0 ? mesh->nf : mesh->nh / 2;
    require((n + 255) / 256 <= kNoAllocMaxBlocks, "too many faces or edges for the allocation-free sum (at most 8388608)");
    char* base = (char*)g_rt.scratch[g_rt.device];
    unsigned* counter = (unsigned*)base;
    double* partials = (double*)(base + kNoAllocCounterBytes);
    if (n > 0) {
        // the last-block counter starts at zero (a fill kernel, not a memset:
        // the call may be captured into a graph)
        launch_fill_ints((int*)counter, 1, 0, nullptr, 0, 0, nullptr, 0, 0, g_rt.stream);
        launch_sum_geometry(mesh->view(), what, partials, counter, out, g_rt.stream);
    } else {
        launch_fill_ints((int*)out, 2, 0, nullptr, 0, 0, nullptr, 0, 0, g_rt.stream);    // +0.0
    }
    check_launch();
    API_END
}

// ---- energy ----
int tm_energy_create(tm_mesh* mesh, const tm_params* params, tm_energy** out)
{
    API_BEGIN
    check_params(*params);
    auto E = std::make_unique<tm_energy>();
    E->params = *params;
    upload_points(*mesh);
    E->nlist = make_nlist(*mesh, E->params);
    // initial_props = properties(mesh) (energy.cpp:27); reference values
    // are not needed for the sums, unit placeholders keep the factors finite
    E->initial = tm_props{1, 1, 1, 0, 0, 0};
    run_pass1(*E, *mesh, 0, nullptr);
    E->initial = props_of(read_globals(*E));
    *out = E.release();
    API_END
}

int tm_energy_destroy(tm_energy* em)
{
    delete em;
    return TM_OK;
}

int tm_energy_get_params(const tm_energy* em, tm_params* out)
{
    *out = em->params;
    return TM_OK;
}

int tm_energy_initial_props(const tm_energy* em, tm_props* out)
{
    *out = em->initial;
    return TM_OK;
}

int tm_energy_set_precision(tm_energy* em, int precision)
{
    API_BEGIN
    require(precision == TM_PREC_FP64 || precision == TM_PREC_MIXED, "unknown precision");
    if (em->precision != precision) em->gen++;
    em->precision = precision;
    API_END
}

int tm_energy_get_precision(const tm_energy* em, int* precision)
{
    *precision = em->precision;
    return TM_OK;
}

int tm_energy_properties(tm_energy* em, tm_mesh* mesh, tm_props* out)
{
    API_BEGIN
    upload_points(*mesh);
    run_pass1(*em, *mesh, em->precision, nullptr);
    *out = props_of(read_globals(*em));
    API_END
}

int tm_energy_vertex_properties(tm_energy* em, tm_mesh* mesh, double* out)
{
    API_BEGIN
    upload_points(*mesh);
    em->vprops.alloc((size_t)mesh->nv * 6);
    run_pass1(*em, *mesh, em->precision, em->vprops.p);
    download_rows(*mesh, em->vprops.p, 6, out);
    sync();
    API_END
}

int tm_energy_energy(tm_energy* em, tm_mesh* mesh, double* out)
{
    API_BEGIN
    upload_points(*mesh);
    run_pass1(*em, *mesh, em->precision, nullptr);
    *out = read_globals(*em).energy;
    API_END
}

int tm_energy_energy_props(tm_energy* em, const tm_props* props, double* out)
{
    *out = energy_from(*em, *props);
    return TM_OK;
}

int tm_energy_gradient(tm_energy* em, tm_mesh* mesh, double* out)
{
    API_BEGIN
    upload_points(*mesh);
    run_pass1(*em, *mesh, em->precision, nullptr);
    run_pass2(*em, *mesh, em->precision);
    download_rows(*mesh, em->grad.p, 3, out);
    sync();
    API_END
}

int tm_energy_energy_gradient(tm_energy* em, tm_mesh* mesh, double* energy, double* grad)
{
    API_BEGIN
    upload_points(*mesh);
    run_pass1(*em, *mesh, em->precision, nullptr);
    run_pass2(*em, *mesh, em->precision);
    Globals g;
    d2h(&g, em->globals.p, sizeof(Globals));
    download_rows(*mesh, em->grad.p, 3, grad);
    sync();
    *energy = g.energy;
    API_END
}

int tm_energy_update_reference_properties(tm_energy* em)
{
    // energy.cpp:47-55
    if (em->params.cont_lambda < 1.0) {
        em->params.cont_lambda += em->params.cont_delta;
        em->gen++;
    }
    return TM_OK;
}

int tm_energy_update_repulsion(tm_energy* em, tm_mesh* mesh)
{
    API_BEGIN
    upload_points(*mesh);
    rebuild_nlist(*em, *mesh);
    API_END
}

int tm_energy_eval_device(tm_energy* em, tm_mesh* mesh, int flags)
{
    API_BEGIN
    ensure_device(*mesh);
    if (flags & TM_EVAL_PROPS) run_pass1(*em, *mesh, em->precision, nullptr);
    if (flags & TM_EVAL_GRAD) run_pass2(*em, *mesh, em->precision);
    API_END
}

const double* tm_energy_device_gradient(tm_energy* em) { return em->grad.p; }
const double* tm_energy_device_globals(tm_energy* em) { return (const double*)em->globals.p; }

int tm_stream(void** out)
{
    API_BEGIN
    ensure_runtime();
    *out = (void*)g_rt.stream;
    API_END
}

int tm_mark_pass1(void* before, void* after)
{
    API_BEGIN
    ensure_runtime();
    g_rt.pass1_marks[0] = (hipEvent_t)before;
    g_rt.pass1_marks[1] = (hipEvent_t)after;
    API_END
}

int tm_sync(void)
{
    API_BEGIN
    ensure_runtime();
    sync();
    API_END
}

// ---- neighbour lists ----
int tm_nlist_create(tm_mesh* mesh, const tm_params* params, tm_nlist** out)
{
    API_BEGIN
    upload_points(*mesh);
    *out = make_nlist(*mesh, *params);
    API_END
}

int tm_nlist_destroy(tm_nlist* nl)
{
    delete nl;
    return TM_OK;
}

int tm_energy_nlist(tm_energy* em, tm_nlist** out)
{
    *out = em->nlist;
    return TM_OK;
}

int tm_nlist_distance_matrix(tm_nlist* nl, tm_mesh* mesh, double dmax, double* d, int* i, int* j, long cap,
                             long* count)
{
    API_BEGIN
    require(nl->nv == mesh->nv, "mesh does not match the neighbour list");
    upload_points(*mesh);
    const int n = mesh->nv;
    nl->dm_counts.alloc(n + 1);
    nl->dm_off.alloc(n + 1);
    HIPCHECK(hipMemsetAsync(nl->dm_counts.p, 0, sizeof(int) * (n + 1), g_rt.stream));
    launch_dm_count(mesh->view(), nl->view(), dmax, nl->dm_counts.p, g_rt.stream);
    check_launch();
    scan(nl->dm_counts.p, nl->dm_off.p, n + 1, nl->scan_tmp);
    int total = 0;
    d2h(&total, nl->dm_off.p + n, sizeof(int));
    sync();
    *count = total;
    if (d && i && j && total > 0) {
        nl->dm_d.alloc(total);
        nl->dm_i.alloc(total);
        nl->dm_j.alloc(total);
        launch_dm_fill(mesh->view(), nl->view(), dmax, nl->dm_off.p, nl->dm_d.p, nl->dm_i.p, nl->dm_j.p,
                       g_rt.stream);
        check_launch();
        long k = std::min<long>(cap, total);
        d2h(d, nl->dm_d.p, sizeof(double) * k);
        d2h(i, nl->dm_i.p, sizeof(int) * k);
        d2h(j, nl->dm_j.p, sizeof(int) * k);
        sync();
    }
    API_END
}

int tm_nlist_point_distances(tm_nlist* nl, tm_mesh* mesh, int pid, double dmax, double* dist3, int* jdx, long cap,
                             long* count)
{
    API_BEGIN
    require(pid >= 0 && pid < mesh->nv, "pid out of range");
    require(nl->nv == mesh->nv, "mesh does not match the neighbour list");
    upload_points(*mesh);
    pid = mesh->iperm[pid];
    // first pass counts, second pass fills a buffer of that size
    nl->pd_count.alloc(1);
    launch_point_distances(mesh->view(), nl->view(), pid, dmax, 0, nl->pd_count.p, nullptr, nullptr, g_rt.stream);
    check_launch();
    int c = 0;
    d2h(&c, nl->pd_count.p, sizeof(int));
    sync();
    *count = c;
    if (dist3 && jdx && c > 0) {
        nl->pd_dist.alloc((size_t)c * 3);
        nl->pd_j.alloc(c);
        launch_point_distances(mesh->view(), nl->view(), pid, dmax, c, nl->pd_count.p, nl->pd_dist.p, nl->pd_j.p,
                               g_rt.stream);
        check_launch();
        long k = std::min<long>(cap, c);
        d2h(dist3, nl->pd_dist.p, sizeof(double) * 3 * k);
        d2h(jdx, nl->pd_j.p, sizeof(int) * k);
        sync();
    }
    API_END
}

// ---- flips ----
// The serial sweep in windows of W proposals (launch_serial_windows).  The
// number of windows a sweep needs depends on how many proposals each
// commits, so they are launched in batches sized from the rate so far, with
// the state read back between batches; surplus windows return at once.
static void serial_windows(tm_energy& E, tm_mesh& M, int nflips, const int* proposals, const double* uniforms,
                           int n_uniforms, uint64_t seed, int injected, int W)
{
    const size_t w = (size_t)W;
    E.sf_state.alloc(1);
    E.sf_int.alloc(w * (2 + 2 + 4 + 4 + 4 * kMaxValence));
    E.sf_dbl.alloc(w * 12);
    E.sf_vstamp.alloc(M.nv);
    E.sf_estamp.alloc(M.nh / 2);
    HIPCHECK(hipMemsetAsync(E.sf_vstamp.p, 0, sizeof(unsigned long long) * M.nv, g_rt.stream));
    HIPCHECK(hipMemsetAsync(E.sf_estamp.p, 0, sizeof(unsigned long long) * (M.nh / 2), g_rt.stream));
    E.sf_fstamp.alloc(M.nf);
    HIPCHECK(hipMemsetAsync(E.sf_fstamp.p, 0, sizeof(unsigned long long) * M.nf, g_rt.stream));
    SerialWindow sw;
    sw.e = E.sf_int.p;
    sw.valid = sw.e + w;
    sw.pv = sw.valid + w;
    sw.fc = sw.pv + 4 * w;
    sw.nbn = sw.fc + 2 * w;
    sw.nb = sw.nbn + 4 * w;
    sw.dO = E.sf_dbl.p;
    sw.dN = sw.dO + 6 * w;
    const EParams p = eparams_of(E);
    const NListView nl = E.nlist->view();
    launch_serial_init(E.globals.p, p, nflips, E.sf_state.p, g_rt.stream);
    int batch = std::max(1, std::min(1024, nflips / W + 1));
    SerialState h{};
    for (;;) {
        launch_serial_windows(M.flip_view(), nl, p, nflips, proposals, uniforms, n_uniforms, seed, injected, W,
                              E.sf_state.p, sw, E.sf_vstamp.p, E.sf_estamp.p, E.sf_fstamp.p, batch, g_rt.stream);
        check_launch();
        d2h(&h, E.sf_state.p, sizeof(SerialState));
        sync();
        if (h.done) break;
        const double per_window = std::max(1.0, (double)h.pos / std::max(h.window, 1));
        batch = std::min(4096, (int)std::ceil((nflips - h.pos) / per_window * 1.1) + 1);
    }
    int res[3] = {h.acc, h.used, h.err};
    h2d(E.flip_result.p, res, sizeof(res));
}

static int flip_impl(tm_energy* em, tm_mesh* mesh, int nflips, const int* proposals, const double* uniforms,
                     int n_uniforms, uint64_t seed, int injected, int* accepted, int* n_used)
{
    API_BEGIN
    upload_points(*mesh);
    refresh_host_topology(*mesh);
    // initial properties and energy (flips.cpp:30-31), FP64
    run_pass1(*em, *mesh, 0, nullptr);
    if (injected && nflips > 0) {
        em->proposals.alloc(nflips);
        h2d(em->proposals.p, proposals, sizeof(int) * nflips);
        em->uniforms.alloc(std::max(n_uniforms, 1));
        if (n_uniforms > 0) h2d(em->uniforms.p, uniforms, sizeof(double) * n_uniforms);
    }
    const int W = em->tune.serial_window;
    if (W == 0) {
        launch_flip_serial(mesh->flip_view(), em->nlist->view(), eparams_of(*em), nflips,
                           injected ? em->proposals.p : nullptr, injected ? em->uniforms.p : nullptr, n_uniforms,
                           seed, injected, em->globals.p, em->flip_result.p, g_rt.stream);
        check_launch();
    } else {
        serial_windows(*em, *mesh, nflips, injected ? em->proposals.p : nullptr, injected ? em->uniforms.p : nullptr,
                       n_uniforms, seed, injected, W);
    }
    HIPCHECK(hipMemsetAsync(mesh->flag.p, 0, sizeof(int), g_rt.stream));
    launch_build_ring(mesh->view(), mesh->ring.p, mesh->meta.p, mesh->flag.p, g_rt.stream);
    check_launch();
    int res[4] = {0, 0, 0, 0};
    d2h(res, em->flip_result.p, sizeof(int) * 3);
    int ovf = 0;
    d2h(&ovf, mesh->flag.p, sizeof(int));
    sync();
    mesh->host_stale = true;
    if (res[0] > 0 || res[2] != 0 || ovf) mesh->topo_version++;   // only accepted flips change topology
    if (res[2] == 1) throw TmError(TM_E_INVALID, "flip proposal out of range");
    if (res[2] == 2) throw TmError(TM_E_INVALID, "uniform stream exhausted");
    if (ovf) throw TmError(TM_E_UNSUPPORTED, "vertex valence exceeds the ring-table width (32)");
    if (accepted) *accepted = res[0];
    if (n_used) *n_used = res[1];
    API_END
}

int tm_flip_parallel(tm_energy* em, tm_mesh* mesh, double flip_ratio, uint64_t seed, int batch, int* accepted)
{
    API_BEGIN
    // flips.cpp:84-85
    if (flip_ratio > 1.0) throw TmError(TM_E_RANGE, "flip_ratio must be <= 1.0");
    const int ne = mesh->nh / 2;
    const int nflips = (int)(ne * flip_ratio);
    // a patch conflicts with proposals on ~100 nearby edges: B = E / 1024
    // keeps the expected fraction of skipped proposals near 10%
    if (batch <= 0) batch = std::max(1, ne / 1024);
    upload_points(*mesh);
    run_pass1(*em, *mesh, 0, nullptr);
    if (!em->pf_owner.p || em->pf_owner.n < (size_t)ne) {
        em->pf_owner.alloc(ne);
        HIPCHECK(hipMemsetAsync(em->pf_owner.p, 0, sizeof(unsigned long long) * ne, g_rt.stream));
    } else {
        HIPCHECK(hipMemsetAsync(em->pf_owner.p, 0, sizeof(unsigned long long) * ne, g_rt.stream));
    }
    em->pf_status.alloc(batch);
    em->pf_dprops.alloc((size_t)12 * batch);    // -old, then new
    em->pf_state.alloc(8);
    if (nflips > 0)
        launch_flip_parallel(mesh->flip_view(), em->nlist->view(), eparams_of(*em), nflips, batch,
                             em->tune.pflip_window, seed,
                             em->globals.p, em->pf_owner.p, em->pf_status.p, em->pf_dprops.p, em->pf_state.p,
                             g_rt.stream);
    check_launch();
    HIPCHECK(hipMemsetAsync(mesh->flag.p, 0, sizeof(int), g_rt.stream));
    launch_build_ring(mesh->view(), mesh->ring.p, mesh->meta.p, mesh->flag.p, g_rt.stream);
    check_launch();
    double st[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (nflips > 0) d2h(st, em->pf_state.p, sizeof(st));
    int ovf = 0;
    d2h(&ovf, mesh->flag.p, sizeof(int));
    sync();
    mesh->host_stale = true;
    if (st[7] > 0 || ovf) mesh->topo_version++;
    if (nflips > 0) std::copy(st, st + 8, em->pf_last);
    if (ovf) throw TmError(TM_E_UNSUPPORTED, "vertex valence exceeds the ring-table width (32)");
    if (accepted) *accepted = (int)st[7];
    API_END
}

int tm_energy_pflip_state(tm_energy* em, double* out)
{
    API_BEGIN
    require(em != nullptr && out != nullptr, "null argument");
    std::copy(em->pf_last, em->pf_last + 8, out);
    API_END
}

int tm_flip_serial(tm_energy* em, tm_mesh* mesh, double flip_ratio, uint64_t seed, int* accepted)
{
    // flip_ratio > 1 is not rejected here, like flips.cpp:24-25
    int nflips = (int)((mesh->nh / 2) * flip_ratio);
    return flip_impl(em, mesh, nflips, nullptr, nullptr, 0, seed, 0, accepted, nullptr);
}

int tm_flip_serial_injected(tm_energy* em, tm_mesh* mesh, int n_proposals, const int* proposals,
                            const double* uniforms, int n_uniforms, int* accepted, int* n_used)
{
    return flip_impl(em, mesh, n_proposals, proposals, uniforms, n_uniforms, 0, 1, accepted, n_used);
}

}    // extern "C"

// ---- device-resident HMC step ----

struct tm_hmc {
    tm_mesh* mesh = nullptr;
    tm_energy* em = nullptr;
    int model = TM_HMC_ENERGY;
    int n3 = 0;
    DevBuf<double> x0, p, a, kpart0, kpart1;
    DevBuf<HmcState> state;
};

extern "C" {

int tm_hmc_create(tm_mesh* mesh, tm_energy* em, int model, tm_hmc** out)
{
    API_BEGIN
    require(mesh != nullptr && out != nullptr, "null argument");
    require(model == TM_HMC_ENERGY || model == TM_HMC_HARMONIC, "unknown HMC energy model");
    require(model != TM_HMC_ENERGY || em != nullptr, "the energy model needs an energy manager");
    ensure_device(*mesh);
    if (em) require(em->nlist && em->nlist->nv == mesh->nv, "mesh does not match the energy manager");
    auto H = std::make_unique<tm_hmc>();
    H->mesh = mesh;
    H->em = model == TM_HMC_ENERGY ? em : nullptr;
    H->model = model;
    H->n3 = 3 * mesh->nv;
    const size_t n3 = std::max(H->n3, 1);
    H->x0.alloc(n3);
    H->p.alloc(n3);
    H->a.alloc(n3);
    H->kpart0.alloc(hmc_blocks(H->n3));
    H->kpart1.alloc(hmc_blocks(H->n3));
    H->state.alloc(1);
    HIPCHECK(hipMemsetAsync(H->state.p, 0, sizeof(HmcState), g_rt.stream));
    sync();
    *out = H.release();
    API_END
}

int tm_hmc_destroy(tm_hmc* hmc)
{
    delete hmc;
    return TM_OK;
}

int tm_hmc_step(tm_hmc* hmc, const tm_hmc_params* prm, double temperature, const double* momenta, double uniform,
                uint64_t seed, uint64_t step)
{
    API_BEGIN
    require(hmc != nullptr && prm != nullptr, "null argument");
    require(prm->num_integration_steps >= 0, "num_integration_steps must be >= 0");
    tm_mesh& M = *hmc->mesh;
    require(3 * M.nv == hmc->n3, "mesh size changed since the HMC handle was created");
    // the device positions are the state; the host array is taken only when
    // it is the newer copy
    if (!M.device_ahead) push_points(M);
    const int n3 = hmc->n3;
    const double mass = prm->mass, dt = prm->time_step;
    const double dt2 = prm->time_step_sq != 0.0 ? prm->time_step_sq : dt * dt;
    const int L = prm->num_integration_steps;
    const int harmonic = hmc->model == TM_HMC_HARMONIC;
    if (momenta && M.nv > 0) {
        h2d(M.stage.p, momenta, sizeof(double) * n3);
        launch_gather_rows(M.stage.p, M.d_perm.p, hmc->p.p, M.nv, 3, g_rt.stream);
        check_launch();
        sync();    // the caller's array may be reused on return
    }
    // p_var = m T (hmc.py:137); sigma = sqrt(p_var) (hmc.py:140)
    launch_hmc_begin(n3, M.x.p, hmc->x0.p, hmc->p.p, M.d_perm.p, momenta ? 0 : 1, seed, step,
                     std::sqrt(mass * temperature), hmc->kpart0.p, g_rt.stream);
    check_launch();
    HmcState* st = hmc->state.p;
    const Globals* globals = nullptr;
    const double* g = M.x.p;    // harmonic: dE/dx = x
    if (!harmonic) {
        tm_energy& E = *hmc->em;
        ensure_energy_buffers(E, M);
        globals = E.globals.p;
        g = E.grad.p;
    }
    // force at x0 (hmc.py:19), then L drift / force / kick steps (hmc.py:20-24)
    for (int i = 0; i <= L; i++) {
        if (!harmonic) {
            run_pass1(*hmc->em, M, hmc->em->precision, nullptr);
            run_pass2(*hmc->em, M, hmc->em->precision);
        }
        int mode = i == 0 ? kHmcFirst : kHmcKick;
        mode |= i < L ? kHmcDrift : kHmcLast;
        launch_hmc_leap(n3, M.x.p, hmc->p.p, hmc->a.p, g, dt, dt2, mass, mode, globals, st, hmc->kpart1.p,
                        g_rt.stream);
        check_launch();
    }
    launch_hmc_accept(n3, hmc->kpart0.p, hmc->kpart1.p, hmc->x0.p, M.x.p, harmonic, globals, mass, temperature,
                      uniform, seed, step, st, g_rt.stream);
    check_launch();
    launch_hmc_restore(n3, M.x.p, hmc->x0.p, st, g_rt.stream);
    check_launch();
    M.device_ahead = true;
    API_END
}

int tm_hmc_result(tm_hmc* hmc, tm_hmc_outcome* out)
{
    API_BEGIN
    require(hmc != nullptr && out != nullptr, "null argument");
    HmcState st;
    d2h(&st, hmc->state.p, sizeof(HmcState));
    sync();
    out->dh = st.dh;
    out->energy0 = st.energy0;
    out->energy1 = st.energy1;
    out->kinetic0 = st.kinetic0;
    out->kinetic1 = st.kinetic1;
    out->uniform = st.uniform;
    out->accepted = st.accepted;
    out->n_accepted = (long long)st.n_accepted;
    out->n_steps = (long long)st.n_steps;
    API_END
}

}    // extern "C"


// ---- replica groups ----
struct tm_replicas {
    int R = 0;
    int mixed = 0;
    std::vector<tm_mesh*> meshes;
    std::vector<tm_energy*> ems;
    // energy passes
    std::vector<PassArgs> h_args;
    DevBuf<PassArgs> d_args;
    std::vector<int> h_block0;
    DevBuf<int> d_block0;
    int total_blocks = 0;
    Tuning tune;
    DevBuf<int> d_block0_rep, d_block0_p2;
    int total_rep = 0, total_p2 = 0;
    // pair cache: per-replica recording flags (device; the reduction clears
    // them), displacement partials, and the prefix tables of the
    // displacement check and of the cached query
    DevBuf<int> force_flags, d_block0_rc;
    DevBuf<double> disp_parts;
    // parallel flip sweeps of the group: per-replica arguments and block
    // prefix tables (rebuilt by every sweep)
    DevBuf<PflipArgs> pf_args;
    DevBuf<int> pf_claim0, pf_eval0, pf_apply0;
    std::vector<int> h_ones;
    int total_disp = 0, total_rc = 0;    // displacement partials, cached-query blocks
    std::vector<unsigned> gens;    // each member's tm_energy::gen when the table was built
    GroupLaunch group() const
    {
        return GroupLaunch{d_args.p,  h_args.data(), R, d_block0.p, total_blocks, d_block0_rep.p, total_rep,
                           d_block0_p2.p, total_p2, tune, mixed, d_block0_rc.p,
                           tune.pair_cache ? total_rc : 0};
    }
    // every replica records at its next evaluation
    void force_all() { h2d(force_flags.p, h_ones.data(), sizeof(int) * R); }
    void force(int r) { h2d(force_flags.p + r, h_ones.data() + r, sizeof(int)); }
    // HMC
    bool hmc_ready = false;
    long long steps_run = 0;    // HMC steps run since the setup
    int L = 0;
    std::vector<HmcArgs> h_hmc;
    DevBuf<HmcArgs> d_hmc;
    std::vector<int> h_hblock0;
    DevBuf<int> d_hblock0;
    int hmc_total = 0;
    std::vector<DevBuf<double>> x0, p, a, kp0, kp1;
    DevBuf<HmcState> states;
    // tempering
    bool has_ladder = false;
    uint64_t pt_seed = 0;
    std::vector<double> h_ladder;
    DevBuf<double> ladder;
    DevBuf<int> owner;
    DevBuf<unsigned long long> pt_round, swaps_acc, swaps_tried;
    // graph of one HMC step
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    void drop_graph()
    {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        graph_exec = nullptr;
        graph = nullptr;
    }
    ~tm_replicas() { drop_graph(); }
};

namespace {

void set_group_shape(tm_replicas& S, const Tuning& t, bool force);

// Rebuild the device table of pass arguments from the replicas' current
// buffers and neighbour lists.  With a ladder, each replica's kappa_b is that
// of the slot it holds on the device.
void refresh_replica_args(tm_replicas& S)
{
    std::vector<int> slot_of(S.R);
    if (S.has_ladder) {
        std::vector<int> owner(S.R);
        d2h(owner.data(), S.owner.p, sizeof(int) * S.R);
        sync();
        for (int s = 0; s < S.R; s++) slot_of[owner[s]] = s;
    }
    S.h_args.resize(S.R);
    S.h_block0.assign(S.R + 1, 0);
    // displacement partials and blocks of the cached query: one per pass-1
    // block (256 vertices) of each replica
    std::vector<int> b0d(S.R + 1, 0), b0c(S.R + 1, 0);
    for (int r = 0; r < S.R; r++) {
        b0d[r + 1] = b0d[r] + pass1_blocks(S.meshes[r]->nv);
        b0c[r + 1] = b0c[r] + pass1_blocks(S.meshes[r]->nv);
    }
    S.total_disp = b0d[S.R];
    S.total_rc = (b0c[S.R] + 7) / 8 * 8;    // a multiple of the 8 XCDs; padding blocks fall past the last replica's vertices
    S.d_block0_rc.alloc(S.R + 1);
    S.disp_parts.alloc(S.total_disp);
    S.force_flags.alloc(S.R);
    S.h_ones.assign(S.R, 1);
    for (int r = 0; r < S.R; r++) {
        tm_mesh& M = *S.meshes[r];
        tm_energy& E = *S.ems[r];
        ensure_device(M);
        ensure_energy_buffers(E, M);
        ensure_masks(*E.nlist, M);
        EParams ep = eparams_of(E);
        if (S.has_ladder) ep.kappa_b = S.h_ladder[slot_of[r]];
        PassBuffers b = buffers(E, M);
        b.pc = pair_cache_of(E, M);
        if (b.pc.x_snap) {
            b.pc.disp_part = S.disp_parts.p + b0d[r];
            b.pc.disp_n = b0d[r + 1] - b0d[r];
            b.pc.dforce = S.force_flags.p + r;
        }
        S.h_args[r] = pass_args(M.view(), E.nlist->view(), ep, b, nullptr, E.grad.p);
        S.h_block0[r + 1] = S.h_block0[r] + S.h_args[r].nblocks;
    }
    S.total_blocks = S.h_block0[S.R];
    S.gens.resize(S.R);
    for (int r = 0; r < S.R; r++) S.gens[r] = S.ems[r]->gen;
    S.d_args.alloc(S.R);
    S.d_block0.alloc(S.R + 1);
    h2d(S.d_args.p, S.h_args.data(), sizeof(PassArgs) * S.R);
    h2d(S.d_block0.p, S.h_block0.data(), sizeof(int) * (S.R + 1));
    h2d(S.d_block0_rc.p, b0c.data(), sizeof(int) * (S.R + 1));
    set_group_shape(S, S.tune, true);
}

// block prefix tables of the repulsion query and pass 2 for tuning t; force:
// every replica records at its next evaluation (new tables, or a pair cache
// switched back on, which must not replay lists it did not keep).  The
// recorded lists do not depend on the launch shape.
void set_group_shape(tm_replicas& S, const Tuning& t, bool force)
{
    std::vector<int> rep(S.R + 1, 0), p2(S.R + 1, 0);
    for (int r = 0; r < S.R; r++) {
        rep[r + 1] = rep[r] + tuned_blocks(S.meshes[r]->nv, t.rep_threads);
        p2[r + 1] = p2[r] + pass2_blocks(S.meshes[r]->nv, t.p2_threads);
    }
    S.tune = t;
    S.tune.rep_fused = 0;     // grouped launches keep the query apart
    if (force) S.force_all();
    S.total_rep = rep[S.R];
    S.total_p2 = p2[S.R];
    S.d_block0_rep.alloc(S.R + 1);
    S.d_block0_p2.alloc(S.R + 1);
    h2d(S.d_block0_rep.p, rep.data(), sizeof(int) * (S.R + 1));
    h2d(S.d_block0_p2.p, p2.data(), sizeof(int) * (S.R + 1));
    sync();
}

// every replica's positions on the device (host arrays win where newer)
void replicas_device_points(tm_replicas& S)
{
    for (tm_mesh* M : S.meshes)
        if (!M->device_ahead) push_points(*M);
}

// The device table current for every member: a member whose manager
// rebuilt its neighbour list (update_repulsion swaps in the spare list),
// advanced its continuation or changed precision since the table was built
// makes the group rebuild the table and drop its captured graph, whose
// launches read the old list.  A precision different from the group's is
// refused.
void replicas_current(tm_replicas& S)
{
    bool stale = false;
    for (int r = 0; r < S.R; r++) {
        if (S.ems[r]->precision != S.mixed)
            throw TmError(TM_E_INVALID, "a replica's precision differs from the group's");
        stale = stale || S.ems[r]->gen != S.gens[r];
    }
    if (!stale) return;
    S.drop_graph();
    refresh_replica_args(S);
}

// exclusion masks current for every replica's topology (the table holds
// the mask buffers, whose addresses do not change)
void replicas_masks(tm_replicas& S)
{
    replicas_current(S);
    for (int r = 0; r < S.R; r++) {
        tm_nlist& L = *S.ems[r]->nlist;
        const unsigned before = L.mask_topo;
        ensure_masks(L, *S.meshes[r]);
        if (L.mask_topo != before) S.force(r);    // new topology: exclusion may have changed
    }
}

void replicas_hmc_step(tm_replicas& S)
{
    launch_hmc_begin_grouped(S.d_hmc.p, S.d_hblock0.p, S.R, S.hmc_total, g_rt.stream);
    for (int i = 0; i <= S.L; i++) {
        launch_pass1_grouped(S.group(), g_rt.stream);
        launch_pass2_grouped(S.group(), g_rt.stream);
        int mode = i == 0 ? kHmcFirst : kHmcKick;
        mode |= i < S.L ? kHmcDrift : kHmcLast;
        launch_hmc_leap_grouped(S.d_hmc.p, S.d_hblock0.p, S.R, S.hmc_total, mode, g_rt.stream);
    }
    launch_hmc_accept_grouped(S.d_hmc.p, S.R, g_rt.stream);
    launch_hmc_restore_grouped(S.d_hmc.p, S.d_hblock0.p, S.R, S.hmc_total, g_rt.stream);
}

}    // namespace

extern "C" {

int tm_replicas_create(tm_mesh* const* meshes, tm_energy* const* ems, int n_replicas, tm_replicas** out)
{
    API_BEGIN
    require(meshes && ems && out && n_replicas > 0, "null argument or no replicas");
    auto S = std::make_unique<tm_replicas>();
    S->R = n_replicas;
    for (int r = 0; r < n_replicas; r++) {
        require(meshes[r] && ems[r], "null replica");
        require(ems[r]->nlist && ems[r]->nlist->nv == meshes[r]->nv, "mesh does not match the energy manager");
        require(ems[r]->precision == ems[0]->precision, "replicas must share one precision");
        S->meshes.push_back(meshes[r]);
        S->ems.push_back(ems[r]);
    }
    S->mixed = ems[0]->precision;
    refresh_replica_args(*S);
    replicas_device_points(*S);
    *out = S.release();
    API_END
}

int tm_replicas_destroy(tm_replicas* rs)
{
    delete rs;
    return TM_OK;
}

int tm_replicas_pflip(tm_replicas* rs, double flip_ratio, const uint64_t* seeds, int* accepted)
{
    API_BEGIN
    require(seeds != nullptr, "null seeds");
    // flips.cpp:84-85
    if (flip_ratio > 1.0) throw TmError(TM_E_RANGE, "flip_ratio must be <= 1.0");
    tm_replicas& S = *rs;
    // each replica's kappa_b is that of the ladder slot it holds now
    std::vector<int> slot_of(S.R);
    if (S.has_ladder) {
        std::vector<int> owner(S.R);
        d2h(owner.data(), S.owner.p, sizeof(int) * S.R);
        sync();
        for (int s = 0; s < S.R; s++) slot_of[owner[s]] = s;
    }
    std::vector<PflipArgs> h(S.R);
    std::vector<int> c0(S.R + 1, 0), e0(S.R + 1, 0), a0(S.R + 1, 0);
    int rounds = 0;
    for (int r = 0; r < S.R; r++) {
        tm_mesh& M = *S.meshes[r];
        tm_energy& E = *S.ems[r];
        // as tm_flip_parallel: the FP64 properties at the current positions
        // start each replica's Metropolis walk
        upload_points(M);
        run_pass1(E, M, 0, nullptr);
        const int ne = M.nh / 2, nflips = (int)(ne * flip_ratio), batch = std::max(1, ne / 1024);
        if (!E.pf_owner.p || E.pf_owner.n < (size_t)ne) E.pf_owner.alloc(ne);
        HIPCHECK(hipMemsetAsync(E.pf_owner.p, 0, sizeof(unsigned long long) * ne, g_rt.stream));
        E.pf_status.alloc(batch);
        E.pf_dprops.alloc((size_t)12 * batch);
        E.pf_state.alloc(8);
        EParams ep = eparams_of(E);
        if (S.has_ladder) ep.kappa_b = S.h_ladder[slot_of[r]];
        h[r] = PflipArgs{M.flip_view(), E.nlist->view(), ep, seeds[r], nflips, batch, E.globals.p, E.pf_owner.p,
                         E.pf_status.p, E.pf_dprops.p, E.pf_dprops.p + 6 * (size_t)batch, E.pf_state.p};
        if (nflips > 0) rounds = std::max(rounds, (nflips + batch - 1) / batch);
        c0[r + 1] = c0[r] + (batch + kClaimPerBlock - 1) / kClaimPerBlock;
        e0[r + 1] = e0[r] + batch;
        a0[r + 1] = a0[r] + (batch + 127) / 128;
    }
    S.pf_args.alloc(S.R);
    S.pf_claim0.alloc(S.R + 1);
    S.pf_eval0.alloc(S.R + 1);
    S.pf_apply0.alloc(S.R + 1);
    h2d(S.pf_args.p, h.data(), sizeof(PflipArgs) * S.R);
    h2d(S.pf_claim0.p, c0.data(), sizeof(int) * (S.R + 1));
    h2d(S.pf_eval0.p, e0.data(), sizeof(int) * (S.R + 1));
    h2d(S.pf_apply0.p, a0.data(), sizeof(int) * (S.R + 1));
    PflipGroup g{S.pf_args.p, S.R, rounds, S.ems[0]->tune.pflip_window, S.pf_claim0.p, c0[S.R],
                 S.pf_eval0.p, e0[S.R], S.pf_apply0.p, a0[S.R]};
    launch_flip_parallel_grouped(g, g_rt.stream);
    check_launch();
    std::vector<double> st(8 * (size_t)S.R, 0.0);
    std::vector<int> ovf(S.R, 0);
    for (int r = 0; r < S.R; r++) {
        tm_mesh& M = *S.meshes[r];
        HIPCHECK(hipMemsetAsync(M.flag.p, 0, sizeof(int), g_rt.stream));
        launch_build_ring(M.view(), M.ring.p, M.meta.p, M.flag.p, g_rt.stream);
        check_launch();
        if (h[r].nflips > 0) d2h(&st[8 * (size_t)r], S.ems[r]->pf_state.p, 8 * sizeof(double));
        d2h(&ovf[r], M.flag.p, sizeof(int));
    }
    sync();
    bool overflow = false;
    for (int r = 0; r < S.R; r++) {
        tm_mesh& M = *S.meshes[r];
        M.host_stale = true;
        if (st[8 * (size_t)r + 7] > 0 || ovf[r]) M.topo_version++;
        if (h[r].nflips > 0) std::copy(&st[8 * (size_t)r], &st[8 * (size_t)r] + 8, S.ems[r]->pf_last);
        overflow = overflow || ovf[r];
        if (accepted) accepted[r] = (int)st[8 * (size_t)r + 7];
    }
    if (overflow) throw TmError(TM_E_UNSUPPORTED, "vertex valence exceeds the ring-table width (32)");
    API_END
}

int tm_replicas_eval(tm_replicas* rs, int flags)
{
    API_BEGIN
    replicas_masks(*rs);
    if (flags & TM_EVAL_PROPS) {
        hipEvent_t marks[2] = {g_rt.pass1_marks[0], g_rt.pass1_marks[1]};
        g_rt.pass1_marks[0] = g_rt.pass1_marks[1] = nullptr;
        launch_pass1_grouped(rs->group(), g_rt.stream, marks);
    }
    if (flags & TM_EVAL_GRAD) launch_pass2_grouped(rs->group(), g_rt.stream);
    check_launch();
    API_END
}

int tm_replicas_globals(tm_replicas* rs, double* out)
{
    API_BEGIN
    std::vector<Globals> g(rs->R);
    for (int r = 0; r < rs->R; r++) d2h(&g[r], rs->ems[r]->globals.p, sizeof(Globals));
    sync();
    for (int r = 0; r < rs->R; r++) {
        for (int k = 0; k < 6; k++) out[7 * r + k] = g[r].props[k];
        out[7 * r + 6] = g[r].energy;
    }
    API_END
}

int tm_replicas_gradient(tm_replicas* rs, int r, double* out)
{
    API_BEGIN
    require(r >= 0 && r < rs->R, "replica index out of range");
    download_rows(*rs->meshes[r], rs->ems[r]->grad.p, 3, out);
    sync();
    API_END
}

int tm_replicas_hmc_setup(tm_replicas* rs, const double* mass, const double* time_step, const double* temperature,
                          const uint64_t* seeds, int num_integration_steps)
{
    API_BEGIN
    require(mass && time_step && temperature && seeds, "null argument");
    require(num_integration_steps >= 0, "num_integration_steps must be >= 0");
    tm_replicas& S = *rs;
    S.drop_graph();
    S.L = num_integration_steps;
    const int R = S.R;
    if (!S.hmc_ready) {
        S.x0.resize(R);
        S.p.resize(R);
        S.a.resize(R);
        S.kp0.resize(R);
        S.kp1.resize(R);
        S.states.alloc(R);
        HIPCHECK(hipMemsetAsync(S.states.p, 0, sizeof(HmcState) * R, g_rt.stream));
    }
    S.h_hmc.resize(R);
    S.h_hblock0.assign(R + 1, 0);
    for (int r = 0; r < R; r++) {
        tm_mesh& M = *S.meshes[r];
        tm_energy& E = *S.ems[r];
        const int n3 = 3 * M.nv;
        const size_t n3a = std::max(n3, 1);
        S.x0[r].alloc(n3a);
        S.p[r].alloc(n3a);
        S.a[r].alloc(n3a);
        S.kp0[r].alloc(hmc_blocks(n3));
        S.kp1[r].alloc(hmc_blocks(n3));
        HmcArgs h{};
        h.n3 = n3;
        h.nblocks = hmc_blocks(n3);
        h.x = M.x.p;
        h.x0 = S.x0[r].p;
        h.p = S.p[r].p;
        h.a = S.a[r].p;
        h.grad = E.grad.p;
        h.globals = E.globals.p;
        h.perm = M.d_perm.p;
        h.st = S.states.p + r;
        h.kpart0 = S.kp0[r].p;
        h.kpart1 = S.kp1[r].p;
        h.mass = mass[r];
        h.dt = time_step[r];
        h.dt2 = time_step[r] * time_step[r];
        h.temperature = temperature[r];
        require(!S.has_ladder || temperature[r] == temperature[0],
                "replica exchange over a kappa_b ladder needs one temperature for all replicas");
        h.seed = seeds[r];
        S.h_hmc[r] = h;
        S.h_hblock0[r + 1] = S.h_hblock0[r] + h.nblocks;
    }
    S.hmc_total = S.h_hblock0[R];
    S.d_hmc.alloc(R);
    S.d_hblock0.alloc(R + 1);
    h2d(S.d_hmc.p, S.h_hmc.data(), sizeof(HmcArgs) * R);
    h2d(S.d_hblock0.p, S.h_hblock0.data(), sizeof(int) * (R + 1));
    sync();
    S.hmc_ready = true;
    API_END
}

int tm_replicas_set_ladder(tm_replicas* rs, const double* ladder, uint64_t seed)
{
    API_BEGIN
    require(ladder != nullptr, "null argument");
    tm_replicas& S = *rs;
    S.h_ladder.assign(ladder, ladder + S.R);
    for (int s = 1; s < S.R; s++) require(S.h_ladder[s] >= S.h_ladder[s - 1], "the ladder must be ascending");
    // the exchange rule min(1, exp(-dk dB / T)) assumes one temperature
    if (S.hmc_ready)
        for (int r = 1; r < S.R; r++)
            require(S.h_hmc[r].temperature == S.h_hmc[0].temperature,
                    "replica exchange over a kappa_b ladder needs one temperature for all replicas");
    S.ladder.alloc(S.R);
    S.owner.alloc(S.R);
    S.pt_round.alloc(1);
    S.swaps_acc.alloc(S.R);
    S.swaps_tried.alloc(S.R);
    std::vector<int> own(S.R);
    for (int s = 0; s < S.R; s++) own[s] = s;
    h2d(S.ladder.p, S.h_ladder.data(), sizeof(double) * S.R);
    h2d(S.owner.p, own.data(), sizeof(int) * S.R);
    HIPCHECK(hipMemsetAsync(S.pt_round.p, 0, sizeof(unsigned long long), g_rt.stream));
    HIPCHECK(hipMemsetAsync(S.swaps_acc.p, 0, sizeof(unsigned long long) * S.R, g_rt.stream));
    HIPCHECK(hipMemsetAsync(S.swaps_tried.p, 0, sizeof(unsigned long long) * S.R, g_rt.stream));
    S.pt_seed = seed;
    S.has_ladder = true;
    refresh_replica_args(S);
    API_END
}

int tm_replicas_swap(tm_replicas* rs)
{
    API_BEGIN
    tm_replicas& S = *rs;
    require(S.has_ladder && S.hmc_ready, "exchange needs a ladder and HMC parameters");
    // the exchange compares the bending of each replica's current state,
    // which an HMC step records
    require(S.steps_run > 0, "exchange needs at least one HMC step first");
    replicas_current(S);
    launch_pt_swap(S.d_args.p, S.d_hmc.p, S.owner.p, S.ladder.p, S.R, S.pt_seed, S.pt_round.p, S.swaps_acc.p,
                   S.swaps_tried.p, g_rt.stream);
    check_launch();
    API_END
}

int tm_replicas_hmc_run(tm_replicas* rs, int n_steps, int swap_every, int use_graph)
{
    API_BEGIN
    tm_replicas& S = *rs;
    require(S.hmc_ready, "tm_replicas_hmc_setup first");
    require(n_steps >= 0, "n_steps must be >= 0");
    require(swap_every <= 0 || S.has_ladder, "exchange needs a ladder");
    replicas_masks(S);
    if (use_graph && !S.graph_exec && n_steps > 0) {
        // the step reads every changing quantity (step counts, kappa_b, list
        // pointers) from device tables, so one capture serves all steps
        HIPCHECK(hipStreamBeginCapture(g_rt.stream, hipStreamCaptureModeThreadLocal));
        replicas_hmc_step(S);
        HIPCHECK(hipStreamEndCapture(g_rt.stream, &S.graph));
        HIPCHECK(hipGraphInstantiate(&S.graph_exec, S.graph, nullptr, nullptr, 0));
    }
    for (int k = 0; k < n_steps; k++) {
        if (use_graph) HIPCHECK(hipGraphLaunch(S.graph_exec, g_rt.stream));
        else replicas_hmc_step(S);
        check_launch();
        if (swap_every > 0 && (k + 1) % swap_every == 0)
            launch_pt_swap(S.d_args.p, S.d_hmc.p, S.owner.p, S.ladder.p, S.R, S.pt_seed, S.pt_round.p,
                           S.swaps_acc.p, S.swaps_tried.p, g_rt.stream);
    }
    check_launch();
    S.steps_run += n_steps;
    for (tm_mesh* M : S.meshes) M->device_ahead = true;
    API_END
}

int tm_replicas_stats(tm_replicas* rs, long long* accepted, long long* steps, int* slot_of, double* props,
                      long long* swaps_accepted, long long* swaps_tried)
{
    API_BEGIN
    tm_replicas& S = *rs;
    const int R = S.R;
    std::vector<HmcState> st(R);
    std::vector<int> own(R);
    std::vector<unsigned long long> sa(R, 0), stt(R, 0);
    if (S.hmc_ready) d2h(st.data(), S.states.p, sizeof(HmcState) * R);
    if (S.has_ladder) {
        d2h(own.data(), S.owner.p, sizeof(int) * R);
        d2h(sa.data(), S.swaps_acc.p, sizeof(unsigned long long) * R);
        d2h(stt.data(), S.swaps_tried.p, sizeof(unsigned long long) * R);
    }
    sync();
    for (int r = 0; r < R; r++) {
        if (accepted) accepted[r] = S.hmc_ready ? (long long)st[r].n_accepted : 0;
        if (steps) steps[r] = S.hmc_ready ? (long long)st[r].n_steps : 0;
        if (props)
            for (int k = 0; k < 6; k++) props[6 * r + k] = S.hmc_ready ? st[r].props[k] : 0.0;
        if (swaps_accepted) swaps_accepted[r] = (long long)sa[r];
        if (swaps_tried) swaps_tried[r] = (long long)stt[r];
    }
    if (slot_of)
        for (int s = 0; s < R; s++) slot_of[S.has_ladder ? own[s] : s] = s;
    API_END
}

int tm_replicas_update_repulsion(tm_replicas* rs)
{
    API_BEGIN
    tm_replicas& S = *rs;
    for (int r = 0; r < S.R; r++) {
        tm_mesh& M = *S.meshes[r];
        tm_energy& E = *S.ems[r];
        upload_points(M);
        rebuild_nlist(E, M);
    }
    refresh_replica_args(S);
    API_END
}

}    // extern "C"

// ---- launch-shape autotuner ----
namespace {

Tuning from_c(const tm_tuning& c)
{
    Tuning t;
    t.rep_threads = c.rep_threads;
    t.rep_batch = c.rep_batch;
    t.p2_threads = c.pass2_threads;
    t.rep_fused = c.rep_fused;
    t.pair_cache = c.pair_cache;
    t.pflip_window = c.pflip_window;
    t.serial_window = c.serial_window;
    return t;
}

tm_tuning to_c(const Tuning& t)
{
    tm_tuning c{};
    c.rep_threads = t.rep_threads;
    c.rep_batch = t.rep_batch;
    c.pass2_threads = t.p2_threads;
    c.rep_fused = t.rep_fused;
    c.pair_cache = t.pair_cache;
    c.pflip_window = t.pflip_window;
    c.serial_window = t.serial_window;
    return c;
}

// mean time of `reps` launches of f after two untimed ones (ms)
template <class F>
double time_launches(F&& f, int reps)
{
    hipEvent_t a, b;
    HIPCHECK(hipEventCreate(&a));
    HIPCHECK(hipEventCreate(&b));
    f();
    f();
    HIPCHECK(hipEventRecord(a, g_rt.stream));
    for (int i = 0; i < reps; i++) f();
    HIPCHECK(hipEventRecord(b, g_rt.stream));
    check_launch();
    HIPCHECK(hipEventSynchronize(b));
    float ms = 0;
    HIPCHECK(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return (double)ms / reps;
}

// Try every shape of the repulsion query and of pass 2 with launch(t) and
// keep the fastest of each; the two kernels are timed separately.  The query
// always has its own launch (rep_fused is kept at 0).
template <class RepF, class P2F>
Tuning search(Tuning base, bool has_rep, RepF&& rep, P2F&& p2, double* best_rep_ms, double* best_p2_ms)
{
    const int ths[4] = {64, 128, 256, 512}, bats[4] = {1, 2, 4, 8};
    Tuning best = base;
    double brep = 1e30, bp2 = 1e30;
    if (has_rep)
        for (int th : ths)
            for (int ba : bats) {
                Tuning t = best;
                t.rep_threads = th;
                t.rep_batch = ba;
                t.rep_fused = 0;
                const double ms = time_launches([&] { rep(t); }, 5);
                if (ms < brep) {
                    brep = ms;
                    best.rep_threads = th;
                    best.rep_batch = ba;
                }
            }
    best.rep_fused = 0;
    for (int th : ths) {
        Tuning t = best;
        t.p2_threads = th;
        const double ms = time_launches([&] { p2(t); }, 5);
        if (ms < bp2) {
            bp2 = ms;
            best.p2_threads = th;
        }
    }
    if (best_rep_ms) *best_rep_ms = has_rep ? brep : 0.0;
    if (best_p2_ms) *best_p2_ms = bp2;
    return best;
}

}    // namespace

extern "C" {

int tm_energy_set_tuning(tm_energy* em, const tm_tuning* t)
{
    API_BEGIN
    require(t != nullptr && tuning_valid(from_c(*t)), "invalid launch shape");
    em->tune = from_c(*t);
    API_END
}

int tm_energy_get_tuning(const tm_energy* em, tm_tuning* out)
{
    *out = to_c(em->tune);
    return TM_OK;
}

int tm_energy_autotune(tm_energy* em, tm_mesh* mesh, tm_tuning* out, double* ms_rep, double* ms_pass2)
{
    API_BEGIN
    tm_energy& E = *em;
    tm_mesh& M = *mesh;
    upload_points(M);
    // pass 2 reads q and the globals; the second pass 1 leaves the pair
    // cache in its steady state (recorded, displacement measured)
    run_pass1(E, M, E.precision, nullptr);
    run_pass1(E, M, E.precision, nullptr);
    const PassArgs a = pass_args(M.view(), E.nlist->view(), eparams_of(E), buffers(E, M), nullptr, E.grad.p);
    if (M.nv > 0)
        E.tune = search(
            E.tune, a.rep_v != nullptr, [&](const Tuning& t) { launch_rep(a, t, g_rt.stream); },
            [&](const Tuning& t) { launch_pass2_kernel(a, E.precision, t, g_rt.stream); }, ms_rep, ms_pass2);
    if (out) *out = to_c(E.tune);
    API_END
}

int tm_replicas_set_tuning(tm_replicas* rs, const tm_tuning* t)
{
    API_BEGIN
    require(t != nullptr && tuning_valid(from_c(*t)), "invalid launch shape");
    rs->drop_graph();
    const bool cache_on = from_c(*t).pair_cache && !rs->tune.pair_cache;
    set_group_shape(*rs, from_c(*t), cache_on);
    API_END
}

int tm_replicas_get_tuning(const tm_replicas* rs, tm_tuning* out)
{
    API_BEGIN
    require(out != nullptr, "null output");
    *out = to_c(rs->tune);
    API_END
}

int tm_replicas_autotune(tm_replicas* rs, tm_tuning* out, double* ms_rep, double* ms_pass2)
{
    API_BEGIN
    tm_replicas& S = *rs;
    replicas_masks(S);
    launch_pass1_grouped(S.group(), g_rt.stream);
    bool has_rep = false;
    for (const PassArgs& a : S.h_args) has_rep = has_rep || a.rep_v;
    // a new shape's tables are uploaded in the untimed warm-up launches
    auto shaped = [&](const Tuning& t) {
        if (t.rep_threads != S.tune.rep_threads || t.rep_batch != S.tune.rep_batch ||
            t.p2_threads != S.tune.p2_threads)
            set_group_shape(S, t, false);
        return S.group();
    };
    const Tuning best = search(
        S.tune, has_rep, [&](const Tuning& t) { launch_rep_grouped(shaped(t), g_rt.stream); },
        [&](const Tuning& t) { launch_pass2_grouped(shaped(t), g_rt.stream); }, ms_rep, ms_pass2);
    S.drop_graph();
    set_group_shape(S, best, false);
    if (out) *out = to_c(best);
    API_END
}

}    // extern "C"

// ---- surrogate force network ----
struct tm_surrogate {
    tm_mesh* mesh = nullptr;
    tm_energy* em = nullptr;
    int nv = 0;
    SgNorm norm{};
    float sigma_F = 1.0f;
    DevBuf<float> theta, grad, m1, m2;
    DevBuf<short> packed;
    SgPacked pk{};
    DevBuf<float> act;    // every activation and work array, carved below
    SgBuffers b{};
    DevBuf<float> geo_buf;
    SgGeom geo{};
    DevBuf<int> twin, err;
    // the feature pass: properties without repulsion, in buffers of its own
    DevBuf<double> q, gprops, vprops, partials, rep_v;
    DevBuf<int> pair_cnt;
    DevBuf<Globals> globals;
    DevBuf<double> neg_force, xsave;
    DevBuf<float> losses;
    std::vector<std::unique_ptr<DevBuf<double>>> snap_x, snap_g;
    int adam_step = 0;
    // plain-gradient fallback of tm_hmc_step_guarded
    int pol_window = 10;         // surrogate-driven outcomes counted together
    double pol_min_acc = 0.5;    // below this acceptance the exact gradient drives
    int pol_probe = 10;          // while it does: one surrogate-driven probe every pol_probe steps
    std::deque<int> recent;      // outcomes (1 accepted) of the last surrogate-driven steps
    bool fallback = false;
    int since_probe = 0;
    bool pending = false;        // the last guarded step was surrogate-driven, its outcome unread
    unsigned checked_topo = ~0u; // topology whose features were last checked
    bool topo_ok = false;        // ... and whether the network can take it (closed, valences 3..10)
    long long n_sg = 0, n_exact = 0, n_switch = 0;
};

namespace {

void sg_alloc(tm_surrogate& S)
{
    const size_t N = (size_t)S.nv;
    const size_t NW = N * kSgW;
    S.theta.alloc(kSgParams);
    S.grad.alloc(kSgParams);
    S.m1.alloc(kSgParams);
    S.m2.alloc(kSgParams);
    HIPCHECK(hipMemsetAsync(S.m1.p, 0, sizeof(float) * kSgParams, g_rt.stream));
    HIPCHECK(hipMemsetAsync(S.m2.p, 0, sizeof(float) * kSgParams, g_rt.stream));
    // packed weights
    const size_t np = (size_t)kSgW * kSgFeatPad + kSgLayers * (4 * kSgW * kSgW + 2 * kSgW * kSgW);
    S.packed.alloc(np);
    short* pp = S.packed.p;
    S.pk.in_t = pp;
    pp += kSgW * kSgFeatPad;
    for (int l = 0; l < kSgLayers; l++) {
        S.pk.u_t[l] = pp; pp += kSgW * kSgW;
        S.pk.wv_t[l] = pp; pp += 2 * kSgW * kSgW;
        S.pk.u_n[l] = pp; pp += kSgW * kSgW;
        S.pk.w_n[l] = pp; pp += kSgW * kSgW;
        S.pk.v_n[l] = pp; pp += kSgW * kSgW;
    }
    // activations
    const size_t nact = N * kSgFeatPad + (2 * (kSgLayers + 1) + 2 * kSgLayers + 4) * NW + 3 * N + (N + 3) / 4 + 1 +
                        sg_partial_floats(S.nv);
    S.act.alloc(nact);
    float* a = S.act.p;
    S.b.f0 = a; a += N * kSgFeatPad;
    for (int l = 0; l <= kSgLayers; l++) { S.b.Z[l] = a; a += NW; S.b.H[l] = a; a += NW; }
    for (int l = 0; l < kSgLayers; l++) { S.b.P[l] = a; a += NW; S.b.M[l] = a; a += NW; }
    S.b.dH = a; a += NW;
    S.b.dZ = a; a += NW;
    S.b.dP = a; a += NW;
    S.b.dM = a; a += NW;
    S.b.ghat = a; a += 3 * N;
    S.b.loss_partial = a; a += (N + 3) / 4;
    S.b.loss = a; a += 1;
    S.b.partial = a;
    S.geo_buf.alloc((size_t)kSgSlots * N * 5 + 3 * N);
    S.geo.eps = S.geo_buf.p;
    S.geo.ehat = S.geo_buf.p + (size_t)kSgSlots * N * 2;
    S.geo.nu = S.geo_buf.p + (size_t)kSgSlots * N * 5;
    S.twin.alloc((size_t)kSgSlots * N);
    S.err.alloc(1);
    S.q.alloc(N);
    S.gprops.alloc(6 * N);
    S.vprops.alloc(6 * N);
    S.partials.alloc((size_t)pass1_blocks(S.nv) * 6);
    S.rep_v.alloc(N);
    S.pair_cnt.alloc(N);
    S.globals.alloc(1);
    S.neg_force.alloc(3 * N);
    S.xsave.alloc(3 * N);
}

// properties without repulsion at the current positions into the
// surrogate's own buffers, then the input features
void sg_feature_pass(tm_surrogate& S)
{
    tm_mesh& M = *S.mesh;
    tm_energy& E = *S.em;
    EParams ep = eparams_of(E);
    ep.rep_lc1 = 0.0;    // no repulsion query: the features do not use it
    S.norm.ref[0] = ep.ref_area;
    S.norm.ref[1] = ep.ref_volume;
    S.norm.ref[2] = ep.ref_curvature;
    PassBuffers pb{};
    pb.q = S.q.p;
    pb.gprops = S.gprops.p;
    pb.pair_cnt = S.pair_cnt.p;
    pb.rep_v = S.rep_v.p;
    pb.partials = S.partials.p;
    pb.globals = S.globals.p;
    pb.list = M.generic.p;
    pb.nlist = 0;
    launch_pass1(M.view(), E.nlist->view(), ep, 1, pb, S.vprops.p, g_rt.stream);
    sg_features(M.view(), S.vprops.p, S.globals.p, S.norm, S.geo, S.b.f0, S.err.p, g_rt.stream);
    check_launch();
}

void sg_check_err(tm_surrogate& S)
{
    int e = 0;
    d2h(&e, S.err.p, sizeof(int));
    sync();
    if (e) {
        HIPCHECK(hipMemsetAsync(S.err.p, 0, sizeof(int), g_rt.stream));
        throw TmError(TM_E_UNSUPPORTED, "surrogate needs a closed mesh with valences 3..10");
    }
}

// the exact gradient at the current positions (energy manager buffers)
void sg_exact_gradient(tm_surrogate& S)
{
    run_pass1(*S.em, *S.mesh, S.em->precision, nullptr);
    run_pass2(*S.em, *S.mesh, S.em->precision);
}

}    // namespace

extern "C" {

int tm_surrogate_create(tm_mesh* mesh, tm_energy* em, uint64_t seed, tm_surrogate** out)
{
    API_BEGIN
    require(mesh && em && out, "null argument");
    require(em->nlist && em->nlist->nv == mesh->nv, "mesh does not match the energy manager");
    tm_mesh& M = *mesh;
    upload_points(M);
    if (M.device_ahead) pull_points(M);    // host copies read below
    refresh_host_topology(M);
    require(M.nv > 0 && M.n_generic == 0, "surrogate needs a closed mesh");
    auto S = std::make_unique<tm_surrogate>();
    S->mesh = mesh;
    S->em = em;
    S->nv = M.nv;
    sg_alloc(*S);
    HIPCHECK(hipMemsetAsync(S->err.p, 0, sizeof(int), g_rt.stream));
    sg_twin(M.view(), S->twin.p, g_rt.stream);
    // mean edge length and raw features of the current state fix the
    // normalisation
    double lsum = 0.0;
    long ne = 0;
    for (int h = 0; h < M.nh; h += 2) {
        const int a = M.perm[M.he_to[h]], b = M.perm[M.he_to[h + 1]];
        const double* pa = M.points + 3 * a;
        const double* pb = M.points + 3 * b;
        lsum += std::sqrt((pa[0] - pb[0]) * (pa[0] - pb[0]) + (pa[1] - pb[1]) * (pa[1] - pb[1]) +
                          (pa[2] - pb[2]) * (pa[2] - pb[2]));
        ne++;
    }
    S->norm.lbar = ne ? lsum / ne : 1.0;
    for (int i = 0; i < kSgFeat; i++) {
        S->norm.mu[i] = 0.0;
        S->norm.sig[i] = 1.0;
    }
    sg_feature_pass(*S);
    sg_check_err(*S);
    std::vector<float> f((size_t)M.nv * kSgFeatPad);
    d2h(f.data(), S->b.f0, sizeof(float) * f.size());
    sync();
    for (int i = 0; i < kSgFeat; i++) {
        double m = 0.0, v = 0.0;
        for (int u = 0; u < M.nv; u++) m += f[(size_t)u * kSgFeatPad + i];
        m /= M.nv;
        for (int u = 0; u < M.nv; u++) {
            const double d = f[(size_t)u * kSgFeatPad + i] - m;
            v += d * d;
        }
        const double sd = std::sqrt(v / M.nv);
        if (i >= 4 && i <= 6) {    // global deviations: already dimensionless
            S->norm.mu[i] = 0.0;
            S->norm.sig[i] = 1.0;
        } else {
            S->norm.mu[i] = m;
            S->norm.sig[i] = sd > 1e-12 * std::fabs(m) + 1e-30 ? sd : 1.0;
        }
    }
    // force scale: rms of the exact force now
    sg_exact_gradient(*S);
    std::vector<double> g((size_t)3 * M.nv);
    d2h(g.data(), em->grad.p, sizeof(double) * g.size());
    sync();
    double ss = 0.0;
    for (double x : g) ss += x * x;
    S->sigma_F = (float)(ss > 0 ? std::sqrt(ss / g.size()) : 1.0);
    // Glorot-uniform weights, small output layer, zero biases
    std::vector<float> th(kSgParams, 0.0f);
    std::mt19937_64 rng(seed);
    auto fill = [&](int off, int fan_in, int fan_out, int count, float gain) {
        const float lim = gain * std::sqrt(6.0f / (float)(fan_in + fan_out));
        std::uniform_real_distribution<float> U(-lim, lim);
        for (int i = 0; i < count; i++) th[off + i] = U(rng);
    };
    fill(kSgOffWin, kSgFeat, kSgW, kSgFeatPad * kSgW, 1.0f);
    for (int i = kSgFeat * kSgW; i < kSgFeatPad * kSgW; i++) th[kSgOffWin + i] = 0.0f;
    for (int l = 0; l < kSgLayers; l++) {
        fill(sg_off_U(l), kSgW, kSgW, kSgW * kSgW, 1.0f);
        fill(sg_off_W(l), 2 * kSgW, kSgW, kSgW * kSgW, 0.5f);
        fill(sg_off_V(l), 2 * kSgW, kSgW, kSgW * kSgW, 0.5f);
        fill(sg_off_E(l), 2, kSgW, 2 * kSgW, 1.0f);
    }
    fill(kSgOffWs, kSgW, 1, kSgW, 0.1f);
    fill(kSgOffWt, kSgW, 1, kSgW, 0.1f);
    h2d(S->theta.p, th.data(), sizeof(float) * kSgParams);
    sg_pack(S->theta.p, S->pk, g_rt.stream);
    check_launch();
    sync();
    *out = S.release();
    API_END
}

int tm_surrogate_destroy(tm_surrogate* sg)
{
    delete sg;
    return TM_OK;
}

int tm_surrogate_params(tm_surrogate* sg, float* out, int* n)
{
    API_BEGIN
    if (n) *n = kSgParams;
    if (out) {
        d2h(out, sg->theta.p, sizeof(float) * kSgParams);
        sync();
    }
    API_END
}

int tm_surrogate_set_params(tm_surrogate* sg, const float* in)
{
    API_BEGIN
    require(in != nullptr, "null argument");
    h2d(sg->theta.p, in, sizeof(float) * kSgParams);
    sg_pack(sg->theta.p, sg->pk, g_rt.stream);
    check_launch();
    sync();
    API_END
}

int tm_surrogate_norm(tm_surrogate* sg, double* mu, double* sig, double* lbar, double* sigma_F)
{
    API_BEGIN
    for (int i = 0; i < kSgFeat; i++) {
        if (mu) mu[i] = sg->norm.mu[i];
        if (sig) sig[i] = sg->norm.sig[i];
    }
    if (lbar) *lbar = sg->norm.lbar;
    if (sigma_F) *sigma_F = sg->sigma_F;
    API_END
}

int tm_surrogate_forces(tm_surrogate* sg, double* forces)
{
    API_BEGIN
    tm_surrogate& S = *sg;
    upload_points(*S.mesh);
    sg_feature_pass(S);
    sg_forward(S.mesh->view(), S.b, S.geo, S.theta.p, S.pk, S.sigma_F, nullptr, S.neg_force.p, g_rt.stream);
    check_launch();
    sg_check_err(S);
    if (forces) {
        std::vector<double> g((size_t)3 * S.nv);
        download_rows(*S.mesh, S.neg_force.p, 3, g.data());
        sync();
        for (size_t i = 0; i < g.size(); i++) forces[i] = -g[i];
    }
    API_END
}

int tm_surrogate_loss_grad(tm_surrogate* sg, double* loss, float* grad)
{
    API_BEGIN
    tm_surrogate& S = *sg;
    upload_points(*S.mesh);
    sg_exact_gradient(S);
    sg_feature_pass(S);
    sg_forward(S.mesh->view(), S.b, S.geo, S.theta.p, S.pk, S.sigma_F, S.em->grad.p, nullptr, g_rt.stream);
    sg_backward(S.mesh->view(), S.b, S.geo, S.theta.p, S.pk, S.twin.p, S.grad.p, g_rt.stream);
    check_launch();
    sg_check_err(S);
    float l = 0.0f;
    d2h(&l, S.b.loss, sizeof(float));
    if (grad) d2h(grad, S.grad.p, sizeof(float) * kSgParams);
    sync();
    if (loss) *loss = l;
    API_END
}

int tm_surrogate_activations(tm_surrogate* sg, int which, float* out)
{
    API_BEGIN
    tm_surrogate& S = *sg;
    require(which >= 0 && which < 1 + (kSgLayers + 1) + kSgLayers, "unknown activation");
    const float* src;
    int w = kSgW;
    if (which == 0) {
        src = S.b.f0;
        w = kSgFeatPad;
    } else if (which <= kSgLayers + 1) {
        src = S.b.H[which - 1];
    } else {
        src = S.b.M[which - kSgLayers - 2];
    }
    std::vector<float> a((size_t)S.nv * w);
    d2h(a.data(), src, sizeof(float) * a.size());
    sync();
    for (int u = 0; u < S.nv; u++) {
        const int i = S.mesh->iperm[u];
        std::memcpy(out + (size_t)u * w, a.data() + (size_t)i * w, sizeof(float) * w);
    }
    API_END
}

int tm_surrogate_add_snapshot(tm_surrogate* sg)
{
    API_BEGIN
    tm_surrogate& S = *sg;
    upload_points(*S.mesh);
    sg_exact_gradient(S);
    auto x = std::make_unique<DevBuf<double>>();
    auto g = std::make_unique<DevBuf<double>>();
    x->alloc((size_t)3 * S.nv);
    g->alloc((size_t)3 * S.nv);
    HIPCHECK(hipMemcpyAsync(x->p, S.mesh->x.p, sizeof(double) * 3 * S.nv, hipMemcpyDeviceToDevice, g_rt.stream));
    HIPCHECK(hipMemcpyAsync(g->p, S.em->grad.p, sizeof(double) * 3 * S.nv, hipMemcpyDeviceToDevice, g_rt.stream));
    S.snap_x.push_back(std::move(x));
    S.snap_g.push_back(std::move(g));
    API_END
}

int tm_surrogate_num_snapshots(tm_surrogate* sg, int* n)
{
    *n = (int)sg->snap_x.size();
    return TM_OK;
}

int tm_surrogate_train(tm_surrogate* sg, int epochs, float lr, uint64_t seed, double* mean_loss)
{
    API_BEGIN
    tm_surrogate& S = *sg;
    const int ns = (int)S.snap_x.size();
    require(ns > 0, "no snapshots recorded");
    require(epochs >= 0, "epochs must be >= 0");
    tm_mesh& M = *S.mesh;
    upload_points(M);
    HIPCHECK(hipMemcpyAsync(S.xsave.p, M.x.p, sizeof(double) * 3 * S.nv, hipMemcpyDeviceToDevice, g_rt.stream));
    S.losses.alloc(std::max(ns, 1));
    std::vector<int> order(ns);
    std::iota(order.begin(), order.end(), 0);
    std::mt19937_64 rng(seed);
    for (int ep = 0; ep < epochs; ep++) {
        std::shuffle(order.begin(), order.end(), rng);
        for (int i = 0; i < ns; i++) {
            const int k = order[i];
            HIPCHECK(hipMemcpyAsync(M.x.p, S.snap_x[k]->p, sizeof(double) * 3 * S.nv, hipMemcpyDeviceToDevice,
                                    g_rt.stream));
            sg_feature_pass(S);
            sg_forward(M.view(), S.b, S.geo, S.theta.p, S.pk, S.sigma_F, S.snap_g[k]->p, nullptr, g_rt.stream);
            sg_backward(M.view(), S.b, S.geo, S.theta.p, S.pk, S.twin.p, S.grad.p, g_rt.stream);
            sg_adam(kSgParams, S.theta.p, S.grad.p, S.m1.p, S.m2.p, lr, ++S.adam_step, g_rt.stream);
            sg_pack(S.theta.p, S.pk, g_rt.stream);
            HIPCHECK(hipMemcpyAsync(S.losses.p + i, S.b.loss, sizeof(float), hipMemcpyDeviceToDevice, g_rt.stream));
            check_launch();
        }
    }
    HIPCHECK(hipMemcpyAsync(M.x.p, S.xsave.p, sizeof(double) * 3 * S.nv, hipMemcpyDeviceToDevice, g_rt.stream));
    std::vector<float> ls(ns);
    d2h(ls.data(), S.losses.p, sizeof(float) * ns);
    sync();
    sg_check_err(S);
    if (mean_loss) {
        double a = 0.0;
        for (float x : ls) a += x;
        *mean_loss = epochs > 0 ? a / ns : 0.0;
    }
    API_END
}

// HMC step with the surrogate's forces along the trajectory and the exact
// energies at its two ends (hmc.py:133-158 with the force replaced)
int tm_hmc_step_surrogate(tm_hmc* hmc, tm_surrogate* sg, const tm_hmc_params* prm, double temperature,
                          const double* momenta, double uniform, uint64_t seed, uint64_t step)
{
    API_BEGIN
    require(hmc && sg && prm, "null argument");
    require(hmc->model == TM_HMC_ENERGY && hmc->em == sg->em && hmc->mesh == sg->mesh,
            "the HMC handle and the surrogate must share mesh and energy manager");
    require(prm->num_integration_steps >= 0, "num_integration_steps must be >= 0");
    tm_mesh& M = *hmc->mesh;
    tm_energy& E = *hmc->em;
    tm_surrogate& S = *sg;
    if (!M.device_ahead) push_points(M);
    const int n3 = hmc->n3;
    const double mass = prm->mass, dt = prm->time_step;
    const double dt2 = prm->time_step_sq != 0.0 ? prm->time_step_sq : dt * dt;
    const int L = prm->num_integration_steps;
    if (momenta) {
        h2d(M.stage.p, momenta, sizeof(double) * n3);
        launch_gather_rows(M.stage.p, M.d_perm.p, hmc->p.p, M.nv, 3, g_rt.stream);
        check_launch();
        sync();
    }
    launch_hmc_begin(n3, M.x.p, hmc->x0.p, hmc->p.p, M.d_perm.p, momenta ? 0 : 1, seed, step,
                     std::sqrt(mass * temperature), hmc->kpart0.p, g_rt.stream);
    HmcState* st = hmc->state.p;
    run_pass1(E, M, E.precision, nullptr);    // exact E(x0), read by the first leap
    for (int i = 0; i <= L; i++) {
        sg_feature_pass(S);
        sg_forward(M.view(), S.b, S.geo, S.theta.p, S.pk, S.sigma_F, nullptr, S.neg_force.p, g_rt.stream);
        int mode = i == 0 ? kHmcFirst : kHmcKick;
        mode |= i < L ? kHmcDrift : kHmcLast;
        launch_hmc_leap(n3, M.x.p, hmc->p.p, hmc->a.p, S.neg_force.p, dt, dt2, mass, mode, E.globals.p, st,
                        hmc->kpart1.p, g_rt.stream);
        check_launch();
    }
    run_pass1(E, M, E.precision, nullptr);    // exact E(xn)
    launch_hmc_accept(n3, hmc->kpart0.p, hmc->kpart1.p, hmc->x0.p, M.x.p, 0, E.globals.p, mass, temperature,
                      uniform, seed, step, st, g_rt.stream);
    launch_hmc_restore(n3, M.x.p, hmc->x0.p, st, g_rt.stream);
    check_launch();
    M.device_ahead = true;
    API_END
}

int tm_surrogate_set_policy(tm_surrogate* sg, int window, double min_acceptance, int probe_every)
{
    API_BEGIN
    require(sg != nullptr, "null argument");
    require(window >= 1 && probe_every >= 1, "window and probe_every must be >= 1");
    require(min_acceptance >= 0.0 && min_acceptance <= 1.0, "min_acceptance must be in [0, 1]");
    sg->pol_window = window;
    sg->pol_min_acc = min_acceptance;
    sg->pol_probe = probe_every;
    sg->recent.clear();
    sg->fallback = false;
    sg->since_probe = 0;
    API_END
}

int tm_surrogate_policy_state(tm_surrogate* sg, int* falling_back, long long* n_surrogate, long long* n_exact,
                              long long* n_switches, double* recent_acceptance)
{
    API_BEGIN
    require(sg != nullptr, "null argument");
    const tm_surrogate& S = *sg;
    if (falling_back) *falling_back = S.fallback;
    if (n_surrogate) *n_surrogate = S.n_sg;
    if (n_exact) *n_exact = S.n_exact;
    if (n_switches) *n_switches = S.n_switch;
    if (recent_acceptance) {
        double a = 0;
        for (int o : S.recent) a += o;
        *recent_acceptance = S.recent.empty() ? -1.0 : a / S.recent.size();
    }
    API_END
}

int tm_hmc_step_guarded(tm_hmc* hmc, tm_surrogate* sg, const tm_hmc_params* prm, double temperature,
                        const double* momenta, double uniform, uint64_t seed, uint64_t step, int* driver)
{
    API_BEGIN
    require(hmc && sg && prm, "null argument");
    tm_surrogate& S = *sg;
    tm_mesh& M = *hmc->mesh;
    // the outcome of the previous surrogate-driven step (waits for its
    // Metropolis kernel), then the switch in either direction
    if (S.pending) {
        HmcState st;
        d2h(&st, hmc->state.p, sizeof(HmcState));
        sync();
        S.pending = false;
        S.recent.push_back(st.accepted ? 1 : 0);
        while ((int)S.recent.size() > S.pol_window) S.recent.pop_front();
        if ((int)S.recent.size() == S.pol_window) {
            double a = 0;
            for (int o : S.recent) a += o;
            a /= S.pol_window;
            if (!S.fallback && a < S.pol_min_acc) {
                S.fallback = true;
                S.since_probe = 0;
                S.recent.clear();
                S.n_switch++;
            } else if (S.fallback && a >= S.pol_min_acc) {
                S.fallback = false;
                S.n_switch++;
            }
        }
    }
    bool use_sg = !S.fallback;
    if (S.fallback && ++S.since_probe >= S.pol_probe) {
        use_sg = true;
        S.since_probe = 0;
    }
    // the network takes closed meshes with valences 3..10 only; the topology
    // changes by flips between steps, so it is checked once per topology
    if (use_sg && S.checked_topo != M.topo_version) {
        if (!M.device_ahead) push_points(M);
        sg_feature_pass(S);
        int e = 0;
        d2h(&e, S.err.p, sizeof(int));
        sync();
        if (e) HIPCHECK(hipMemsetAsync(S.err.p, 0, sizeof(int), g_rt.stream));
        S.topo_ok = e == 0;
        S.checked_topo = M.topo_version;
    }
    use_sg = use_sg && S.topo_ok;
    int rc;
    if (use_sg) {
        rc = tm_hmc_step_surrogate(hmc, sg, prm, temperature, momenta, uniform, seed, step);
        S.pending = rc == TM_OK;
        S.n_sg++;
    } else {
        rc = tm_hmc_step(hmc, prm, temperature, momenta, uniform, seed, step);
        S.n_exact++;
    }
    if (rc != TM_OK) throw TmError(rc, tm_last_error());
    if (driver) *driver = use_sg ? 1 : 0;
    API_END
}

}    // extern "C"
