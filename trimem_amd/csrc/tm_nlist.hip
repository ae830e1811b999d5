// tm_nlist.hip -- cell-list binning and neighbour queries on the GPU.
//
// Reference: src/nlists/cell_list.h:28-313 and verlet_list.h:14-103.
//
// Build (cell list): a box reduction reproducing the reference's partial
// initialisation (cell_list.h:46-47: only component 0 of each bound starts at
// +max / -1, components 1 and 2 start at 0), the per-dimension cell size
// r_list = dim / int(dim / rlist) (cell_list.h:59-66, computed on the host in
// FP64 from the reduced box), flat cell ids sum_k int((p_k - min_k) / r_k)
// * stride_k (cell_list.h:70-79), then a counting sort: histogram, exclusive
// scan, scatter, and a per-cell sort by vertex id so that the order inside a
// cell is ascending like the reference's push_back order.  The reference keys
// its cells in a map (cell_list.h:31), so only occupied cells cost memory;
// here the table is dense over the occupied sub-grid, the global cells lo ..
// hi that hold a vertex at build time, which is the mesh's own bounding box
// in cells whatever the reference's box (anchored at 0 in y and z) spans.
// Stencils clipped to the sub-grid visit the same vertices as stencils
// clipped to the global grid: every cell outside the sub-grid is empty.
//
// Queries use the current positions with the cells frozen at build time,
// exactly as the reference does between two update_repulsion calls.  All
// distance tests are IEEE FP64 sqrt(dx*dx + dy*dy + dz*dz) (the library is
// built with -ffp-contract=off), compared with < dmax (point_distances,
// cell_list.h:208) or <= rlist (distance_matrix, cell_list.h:294).
#include "tm_internal.h"
#include "tm_om.h"

#include <cfloat>

namespace tmh {

// ----------------------------------------------------------------- box ----
// Bounding box of the vertices, as kBoxBlocks per-block partial minima and
// maxima (part[6 b + k], min in k < 3, max in k >= 3; +DBL_MAX and -DBL_MAX
// for a block without vertices).  The host, which reads the box back anyway
// to size the grid, combines them (min and max are exact in any order) and
// applies the reference's initial values (cell_list.h:46-47); a last-block
// combine would need a device-scope fence in every block, which costs more
// than the whole pass.
constexpr int kBoxBlocks = 64;
__global__ void __launch_bounds__(256) k_box(const double* __restrict__ x, int n, double* __restrict__ part)
{
    __shared__ double s[6][4];
    double mn[3] = {DBL_MAX, DBL_MAX, DBL_MAX};
    double mx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    // kBoxBatch points per thread and step, all their loads in flight before
    // the first comparison
    constexpr int kBoxBatch = 8;
    const int stride = gridDim.x * blockDim.x;
    for (int i0 = blockIdx.x * blockDim.x + threadIdx.x; i0 < n; i0 += kBoxBatch * stride) {
        double v[kBoxBatch][3];
#pragma unroll
        for (int u = 0; u < kBoxBatch; u++) {
            const int i = i0 + u * stride;
#pragma unroll
            for (int k = 0; k < 3; k++) v[u][k] = i < n ? x[3 * i + k] : __builtin_nan("");    // neither test takes NaN
        }
#pragma unroll
        for (int u = 0; u < kBoxBatch; u++)
#pragma unroll
            for (int k = 0; k < 3; k++) {
                if (v[u][k] < mn[k]) mn[k] = v[u][k];
                if (v[u][k] > mx[k]) mx[k] = v[u][k];
            }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int k = 0; k < 3; k++) {
        for (int off = 32; off > 0; off >>= 1) {
            mn[k] = fmin(mn[k], __shfl_xor(mn[k], off, 64));
            mx[k] = fmax(mx[k], __shfl_xor(mx[k], off, 64));
        }
    }
    if (lane == 0) {
        for (int k = 0; k < 3; k++) {
            s[k][wave] = mn[k];
            s[3 + k][wave] = mx[k];
        }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        int k = threadIdx.x;
        double r = s[k][0];
        for (int w = 1; w < 4; w++) r = k < 3 ? fmin(r, s[k][w]) : fmax(r, s[k][w]);
        part[blockIdx.x * 6 + k] = r;
    }
}

int box_scratch_doubles() { return kBoxBlocks * 6; }

// Up to three int ranges set to a value each, in one launch (the list build
// issues these fills back to back after its one host round trip).
__global__ void k_fill_ints(int* a, int na, int va, int* b, int nb, int vb, int* c, int nc, int vc)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < na) a[i] = va;
    if (i < nb) b[i] = vb;
    if (i < nc) c[i] = vc;
}

void launch_fill_ints(int* a, int na, int va, int* b, int nb, int vb, int* c, int nc, int vc, hipStream_t s)
{
    const int n = std::max(na, std::max(nb, nc));
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill_ints, dim3((n + 255) / 256), dim3(256), 0, s, a, na, va, b, nb, vb, c, nc, vc);
}

void launch_box(const double* x, int n, double* part, hipStream_t s)
{
    hipLaunchKernelGGL(k_box, dim3(kBoxBlocks), dim3(256), 0, s, x, n, part);
}

// ------------------------------------------------------------- binning ----
// Build-time cell of vertex v in the occupied sub-grid.  The coordinate is
// clamped into the sub-grid: the reference's box epsilon (1e-6) vanishes in
// rounding for coordinates beyond ~1e9, where int((max - min) / r) can reach
// the grid size.
__device__ __forceinline__ int flat_cell(const NListView& nl, const double* x, int v)
{
    int id = 0;
    for (int k = 0; k < 3; k++) {
        int c = cell_coord(nl, x[3 * v + k], k);
        c = min(max(c, 0), nl.shape[k] - 1);
        id += c * nl.strides[k];
    }
    return id;
}

__global__ void k_cell_ids(const double* __restrict__ x, int n, NListView nl, int* __restrict__ vert_cell,
                           int* __restrict__ counts)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    int c = flat_cell(nl, x, v);
    vert_cell[v] = c;
    atomicAdd(&counts[c], 1);
}

void launch_cell_ids(const double* x, int n, const NListView& nl, int* vert_cell, int* counts, hipStream_t s)
{
    hipLaunchKernelGGL(k_cell_ids, dim3((n + 255) / 256), dim3(256), 0, s, x, n, nl, vert_cell, counts);
}

// ------------------------------------------------ stencil candidate lists
// The 27-cell stencil of cell (c0, c1, c2) as 9 rows of x-consecutive cells,
// clipped to the grid like the query of cell_list.h:133-219; f(start, len)
// gets each row's range in cell_verts.
template <class F>
__device__ __forceinline__ void for_stencil_rows(const NListView& nl, const int c[3], F&& f)
{
    const int i0 = max(c[0] - 1, 0), i1 = min(c[0] + 1, nl.shape[0] - 1);
    if (i0 > i1) return;
    for (int r = 0; r < 9; r++) {
        const int kk = c[2] + r / 3 - 1, jj = c[1] + r % 3 - 1;
        if (kk < 0 || kk >= nl.shape[2] || jj < 0 || jj >= nl.shape[1]) continue;
        const int base = jj * nl.strides[1] + kk * nl.strides[2];
        const int s0 = nl.cell_start[base + i0];
        f(s0, nl.cell_start[base + i1 + 1] - s0);
    }
}

__device__ __forceinline__ bool cand_leader(const NListView& nl, int slot, int c[3])
{
    const int cf = nl.vert_cell[nl.cell_verts[slot]];
    if (nl.cell_start[cf] != slot) return false;
    c[2] = cf / nl.strides[2];
    c[1] = (cf - c[2] * nl.strides[2]) / nl.strides[1];
    c[0] = cf - c[2] * nl.strides[2] - c[1] * nl.strides[1];
    return true;
}

// count[slot] = stencil size for the first slot of each occupied cell, else 0
__global__ void k_cand_count(NListView nl, int n, int* __restrict__ count)
{
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= n) return;
    int c[3], total = 0;
    if (cand_leader(nl, slot, c)) for_stencil_rows(nl, c, [&](int, int len) { total += len; });
    count[slot] = total;
}

__global__ void k_cand_fill(NListView nl, int n, const int* __restrict__ cand_start, int* __restrict__ cand)
{
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= n) return;
    int c[3];
    if (!cand_leader(nl, slot, c)) return;
    int o = cand_start[slot];
    for_stencil_rows(nl, c, [&](int s0, int len) {
        for (int q = 0; q < len; q++) cand[o++] = nl.cell_verts[s0 + q];
    });
}

// vcand[v] = list range of the build-time cell of v
__global__ void k_cand_vrange(NListView nl, int n, const int* __restrict__ cand_start, int2* __restrict__ vcand)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const int s = nl.cell_start[nl.vert_cell[v]];
    vcand[v] = make_int2(cand_start[s], cand_start[s + 1]);
}

void launch_cand_vrange(const NListView& nl, int n, const int* cand_start, int2* vcand, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_cand_vrange, dim3((n + 255) / 256), dim3(256), 0, s, nl, n, cand_start, vcand);
}

void launch_cand_count(const NListView& nl, int n, int* count, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_cand_count, dim3((n + 255) / 256), dim3(256), 0, s, nl, n, count);
}

void launch_cand_fill(const NListView& nl, int n, const int* cand_start, int* cand, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_cand_fill, dim3((n + 255) / 256), dim3(256), 0, s, nl, n, cand_start, cand);
}

__global__ void k_cell_scatter(const int* __restrict__ vert_cell, int n, const int* __restrict__ cell_start,
                               int* __restrict__ fill, int* __restrict__ cell_verts)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    int c = vert_cell[v];
    int pos = atomicAdd(&fill[c], 1);
    cell_verts[cell_start[c] + pos] = v;
}

void launch_cell_scatter(const int* vert_cell, int n, const int* cell_start, int* fill, int* cell_verts,
                         hipStream_t s)
{
    hipLaunchKernelGGL(k_cell_scatter, dim3((n + 255) / 256), dim3(256), 0, s, vert_cell, n, cell_start, fill,
                       cell_verts);
}

// The thread sitting at the start of a cell's segment sorts the segment by
// user vertex id (cells hold a handful of vertices), the order in which the
// reference fills its cells.
__global__ void k_cell_sort_segments(const int* __restrict__ vert_cell, int n, const int* __restrict__ cell_start,
                                     const int* __restrict__ perm, int* __restrict__ cell_verts)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int c = vert_cell[cell_verts[i]];
    if (cell_start[c] != i) return;
    int e = cell_start[c + 1];
    for (int a = i + 1; a < e; a++) {
        int key = cell_verts[a];
        int kp = perm[key];
        int b = a - 1;
        while (b >= i && perm[cell_verts[b]] > kp) {
            cell_verts[b + 1] = cell_verts[b];
            b--;
        }
        cell_verts[b + 1] = key;
    }
}

void launch_cell_sort_segments(const int* vert_cell, int n, const int* cell_start, const int* perm, int* cell_verts,
                               hipStream_t s)
{
    hipLaunchKernelGGL(k_cell_sort_segments, dim3((n + 255) / 256), dim3(256), 0, s, vert_cell, n, cell_start, perm,
                       cell_verts);
}

// ---------------------------------------------------- exclusive scan ------
// Three phases: block-local scans of 4096 elements with block totals, a scan
// of the totals by one block, and the add-back.
constexpr int kScanThreads = 1024;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanThreads * kScanItems;

int scan_block_sums_size(int n) { return (n + kScanTile - 1) / kScanTile + 1; }

__device__ __forceinline__ int block_exclusive_scan(int v, int* s_warp, int* total)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int x = v;
    for (int off = 1; off < 64; off <<= 1) {
        int y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) s_warp[wave] = x;
    __syncthreads();
    if (wave == 0) {
        int nw = blockDim.x >> 6;
        int w = lane < nw ? s_warp[lane] : 0;
        for (int off = 1; off < 64; off <<= 1) {
            int y = __shfl_up(w, off, 64);
            if (lane >= off) w += y;
        }
        if (lane < nw) s_warp[lane] = w;
    }
    __syncthreads();
    int base = wave > 0 ? s_warp[wave - 1] : 0;
    *total = s_warp[(blockDim.x >> 6) - 1];
    __syncthreads();
    return base + x - v;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_tiles(const int* __restrict__ in, int* __restrict__ out,
                                                             int n, int* __restrict__ sums)
{
    __shared__ int s_warp[16];
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    int loc[kScanItems];
    int acc = 0;
    for (int k = 0; k < kScanItems; k++) {
        int v = (base + k < n) ? in[base + k] : 0;
        loc[k] = acc;
        acc += v;
    }
    int total;
    int off = block_exclusive_scan(acc, s_warp, &total);
    for (int k = 0; k < kScanItems; k++)
        if (base + k < n) out[base + k] = off + loc[k];
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_sums(int* __restrict__ sums, int nb)
{
    __shared__ int s_warp[16];
    int carry = 0;
    for (int b0 = 0; b0 < nb; b0 += kScanThreads) {
        int i = b0 + threadIdx.x;
        int v = i < nb ? sums[i] : 0;
        int total;
        int ex = block_exclusive_scan(v, s_warp, &total);
        if (i < nb) sums[i] = carry + ex;
        carry += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) sums[nb] = carry;
}

__global__ void k_scan_add(int* __restrict__ out, int n, const int* __restrict__ sums)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += sums[i / kScanTile];
}

void launch_exclusive_scan(const int* in, int* out, int n, int* block_sums, hipStream_t s)
{
    int nb = (n + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(k_scan_tiles, dim3(nb), dim3(kScanThreads), 0, s, in, out, n, block_sums);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanThreads), 0, s, block_sums, nb);
    hipLaunchKernelGGL(k_scan_add, dim3((n + 255) / 256), dim3(256), 0, s, out, n, block_sums);
}

// ------------------------------------------------------ distance matrix ---
__device__ __forceinline__ double dist_ab(const double* x, int a, int b)
{
    // cell_list.h:11-25 squ_distance / distance
    double d0 = x[3 * a] - x[3 * b], d1 = x[3 * a + 1] - x[3 * b + 1], d2 = x[3 * a + 2] - x[3 * b + 2];
    double s = 0.0;
    s += d0 * d0;
    s += d1 * d1;
    s += d2 * d2;
    return sqrt(s);
}

// Visit the pairs a emits: for the cell list, b in the 27 cells around a's
// build-time cell with (cell(b) > cell(a)) or (same cell and b > a), not
// excluded, distance <= dmax -- each unordered pair of the reference's
// cell-pair enumeration exactly once (cell_list.h:221-312); for the Verlet
// list, b in adj(a) with distance < dmax (verlet_list.h:39-72).
template <class F>
__device__ __forceinline__ void dm_visit(const MeshView& m, const NListView& nl, double dmax, int a, F&& f)
{
    if (nl.kind == 0) {
        const Topo t = topo_of(m);
        int ca = nl.vert_cell[a];
        int kc = ca / nl.strides[2];
        int rem = ca % nl.strides[2];
        int jc = rem / nl.strides[1];
        int ic = rem % nl.strides[1];
        for (int di = -1; di < 2; di++) {
            int ii = ic + di;
            if (ii < 0 || ii >= nl.shape[0]) continue;
            for (int dj = -1; dj < 2; dj++) {
                int jj = jc + dj;
                if (jj < 0 || jj >= nl.shape[1]) continue;
                for (int dk = -1; dk < 2; dk++) {
                    int kk = kc + dk;
                    if (kk < 0 || kk >= nl.shape[2]) continue;
                    int oc = ii * nl.strides[0] + jj * nl.strides[1] + kk * nl.strides[2];
                    if (oc < ca) continue;
                    for (int q = nl.cell_start[oc]; q < nl.cell_start[oc + 1]; q++) {
                        int b = nl.cell_verts[q];
                        if (oc == ca && b <= a) continue;
                        double d = dist_ab(m.x, a, b);
                        if (d <= dmax && !om_excluded(t, a, b, nl.exclusion)) f(b, d);
                    }
                }
            }
        }
    } else if (nl.kind == 1) {
        for (int q = nl.adj_start[a]; q < nl.adj_start[a + 1]; q++) {
            int b = nl.adj[q];
            double d = dist_ab(m.x, a, b);
            if (d < dmax) f(b, d);
        }
    }
}

__global__ void k_dm_count(MeshView m, NListView nl, double dmax, int* __restrict__ counts)
{
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= m.nv) return;
    int c = 0;
    dm_visit(m, nl, dmax, a, [&](int, double) { c++; });
    counts[a] = c;
}

__global__ void k_dm_fill(MeshView m, NListView nl, double dmax, const int* __restrict__ offsets, double* d,
                          int* ii, int* jj)
{
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= m.nv) return;
    int o = offsets[a];
    const bool ordered = nl.kind == 0;
    const int ua = nl.perm[a];
    dm_visit(m, nl, dmax, a, [&](int b, double dd) {
        const int ub = nl.perm[b];
        d[o] = dd;
        if (ordered) {
            ii[o] = ua < ub ? ua : ub;
            jj[o] = ua < ub ? ub : ua;
        } else {
            ii[o] = ua;
            jj[o] = ub;
        }
        o++;
    });
}

void launch_dm_count(const MeshView& m, const NListView& nl, double dmax, int* counts, hipStream_t s)
{
    hipLaunchKernelGGL(k_dm_count, dim3((m.nv + 127) / 128), dim3(128), 0, s, m, nl, dmax, counts);
}

void launch_dm_fill(const MeshView& m, const NListView& nl, double dmax, const int* offsets, double* d, int* i,
                    int* j, hipStream_t s)
{
    hipLaunchKernelGGL(k_dm_fill, dim3((m.nv + 127) / 128), dim3(128), 0, s, m, nl, dmax, offsets, d, i, j);
}

// ------------------------------------------------------ point distances ---
// One vertex, in the reference's output order: stencil loops i (x), j (y),
// k (z) around the cell of the current position (cell_list.h:147-213), then
// ascending ids inside a cell; Verlet: adjacency order.
__global__ void k_point_distances(MeshView m, NListView nl, int pid, double dmax, int cap, int* count,
                                  double* dist3, int* jdx)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const Topo t = topo_of(m);
    const d3 p = t.P(pid);
    int c = 0;
    auto emit = [&](int j) {
        d3 d = p - t.P(j);
        if (norm(d) < dmax) {
            if (c < cap) {
                dist3[3 * c] = d.x;
                dist3[3 * c + 1] = d.y;
                dist3[3 * c + 2] = d.z;
                jdx[c] = nl.perm[j];
            }
            c++;
        }
    };
    if (nl.kind == 0) {
        int co[3];
        cell_coords(nl, p, co);
        for (int i = -1; i < 2; i++) {
            int ii = co[0] + i;
            if (ii < 0 || ii >= nl.shape[0]) continue;
            for (int j = -1; j < 2; j++) {
                int jj = co[1] + j;
                if (jj < 0 || jj >= nl.shape[1]) continue;
                for (int k = -1; k < 2; k++) {
                    int kk = co[2] + k;
                    if (kk < 0 || kk >= nl.shape[2]) continue;
                    int oc = ii * nl.strides[0] + jj * nl.strides[1] + kk * nl.strides[2];
                    for (int q = nl.cell_start[oc]; q < nl.cell_start[oc + 1]; q++) {
                        int jid = nl.cell_verts[q];
                        if (om_excluded(t, pid, jid, nl.exclusion)) continue;
                        emit(jid);
                    }
                }
            }
        }
    } else if (nl.kind == 1) {
        for (int q = nl.adj_start[pid]; q < nl.adj_start[pid + 1]; q++) emit(nl.adj[q]);
    }
    *count = c;
}

void launch_point_distances(const MeshView& m, const NListView& nl, int pid, double dmax, int cap, int* count,
                            double* dist3, int* jdx, hipStream_t s)
{
    hipLaunchKernelGGL(k_point_distances, dim3(1), dim3(64), 0, s, m, nl, pid, dmax, cap, count, dist3, jdx);
}

// --------------------------------------------------------- verlet build ---
// VerletList ctor (verlet_list.h:18-35): the symmetric adjacency of the cell
// list's distance_matrix(rlist) pairs.  Per vertex a: every b in the stencil
// of a's cell (either side of the cell ordering), not excluded, distance <=
// rlist; stored ascending.
template <class F>
__device__ __forceinline__ void verlet_visit(const MeshView& m, const NListView& nl, double rlist, int a, F&& f)
{
    const Topo t = topo_of(m);
    int ca = nl.vert_cell[a];
    int kc = ca / nl.strides[2];
    int rem = ca % nl.strides[2];
    int jc = rem / nl.strides[1];
    int ic = rem % nl.strides[1];
    for (int dk = -1; dk < 2; dk++) {
        int kk = kc + dk;
        if (kk < 0 || kk >= nl.shape[2]) continue;
        for (int dj = -1; dj < 2; dj++) {
            int jj = jc + dj;
            if (jj < 0 || jj >= nl.shape[1]) continue;
            int i0 = max(ic - 1, 0), i1 = min(ic + 1, nl.shape[0] - 1);
            int base = jj * nl.strides[1] + kk * nl.strides[2];
            for (int q = nl.cell_start[base + i0]; q < nl.cell_start[base + i1 + 1]; q++) {
                int b = nl.cell_verts[q];
                if (b == a) continue;
                if (dist_ab(m.x, a, b) <= rlist && !om_excluded(t, a, b, nl.exclusion)) f(b);
            }
        }
    }
}

__global__ void k_verlet_count(MeshView m, NListView nl, double rlist, int* __restrict__ counts)
{
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= m.nv) return;
    int c = 0;
    verlet_visit(m, nl, rlist, a, [&](int) { c++; });
    counts[a] = c;
}

__global__ void k_verlet_fill(MeshView m, NListView nl, double rlist, const int* __restrict__ offsets,
                              int* __restrict__ adj)
{
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= m.nv) return;
    int s = offsets[a];
    int o = s;
    verlet_visit(m, nl, rlist, a, [&](int b) { adj[o++] = b; });
    for (int p = s + 1; p < o; p++) {
        int key = adj[p];
        int q = p - 1;
        while (q >= s && adj[q] > key) {
            adj[q + 1] = adj[q];
            q--;
        }
        adj[q + 1] = key;
    }
}

void launch_verlet_count(const MeshView& m, const NListView& cl, double rlist, int* counts, hipStream_t s)
{
    hipLaunchKernelGGL(k_verlet_count, dim3((m.nv + 127) / 128), dim3(128), 0, s, m, cl, rlist, counts);
}

void launch_verlet_fill(const MeshView& m, const NListView& cl, double rlist, const int* offsets, int* adj,
                        hipStream_t s)
{
    hipLaunchKernelGGL(k_verlet_fill, dim3((m.nv + 127) / 128), dim3(128), 0, s, m, cl, rlist, offsets, adj);
}

}    // namespace tmh
