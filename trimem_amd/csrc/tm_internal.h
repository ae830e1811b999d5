// tm_internal.h -- launchers shared between the library's translation units.
#pragma once

#include "tm_common.h"

namespace tmh {

// energy passes (tm_energy.hip)
struct PassBuffers {
    double* q;          // [N] curvature / area per vertex
    double* gprops;     // [6N] properties of listed (boundary) vertices
    int* pair_cnt;      // [N] repulsion pairs per vertex (k_rep)
    double* rep_v;      // [N] repulsion property per vertex (k_rep)
    double* partials;   // [6 * pass1 blocks]
    unsigned* counter;  // last-block counter of pass 1 (returns to 0)
    Globals* globals;
    const int* list;    // boundary / degenerate vertices
    int nlist;
};
int pass1_blocks(int nv);
void launch_pass1(const MeshView& m, const NListView& nl, const EParams& p, int mixed, const PassBuffers& b,
                  double* vprops, hipStream_t s);
void launch_pass2(const MeshView& m, const NListView& nl, const EParams& p, int mixed, const PassBuffers& b,
                  double* grad, hipStream_t s);
void launch_halfedge_geometry(const MeshView& m, int h, double* out, hipStream_t s);
void launch_sum_geometry(const MeshView& m, int what, double* partials, unsigned* counter, double* result,
                         hipStream_t s);

// dst[i*w + c] = src[idx[i]*w + c] for i < n, c < w (tm_flips.hip)
void launch_gather_rows(const double* src, const int* idx, double* dst, int n, int w, hipStream_t s);

// topology helpers (tm_flips.hip)
// Rebuild ring table and meta from the halfedge arrays; sets *overflow to 1
// if a vertex has more than kMaxValence neighbours.
void launch_build_ring(const MeshView& m, int* ring, int* meta, int* overflow, hipStream_t s);

// neighbour lists (tm_nlist.hip)
int box_scratch_doubles();
void launch_box(const double* x, int n, double* scratch, unsigned* counter, double* box6, hipStream_t s);
void launch_cell_ids(const double* x, int n, const NListView& nl, int* vert_cell, int* counts, hipStream_t s);
void launch_exclusive_scan(const int* in, int* out, int n, int* block_sums, hipStream_t s);
int scan_block_sums_size(int n);
void launch_cell_scatter(const int* vert_cell, int n, const int* cell_start, int* fill, int* cell_verts,
                         hipStream_t s);
void launch_cell_sort_segments(const int* vert_cell, int n, const int* cell_start, const int* perm, int* cell_verts,
                               hipStream_t s);
// per occupied cell, the vertices of its 27-cell stencil (count, then fill
// after an exclusive scan of the counts)
void launch_cand_count(const NListView& nl, int n, int* count, hipStream_t s);
void launch_cand_fill(const NListView& nl, int n, const int* cand_start, int* cand, hipStream_t s);
void launch_cand_vrange(const NListView& nl, int n, const int* cand_start, int2* vcand, hipStream_t s);
// distance_matrix (cell list or verlet) in two phases: counts then fill
void launch_dm_count(const MeshView& m, const NListView& nl, double dmax, int* counts, hipStream_t s);
void launch_dm_fill(const MeshView& m, const NListView& nl, double dmax, const int* offsets, double* d, int* i,
                    int* j, hipStream_t s);
// point_distances of one vertex in the reference's stencil order
void launch_point_distances(const MeshView& m, const NListView& nl, int pid, double dmax, int cap, int* count,
                            double* dist3, int* jdx, hipStream_t s);
// verlet adjacency from a cell list: counts, then fill (ascending ids)
void launch_verlet_count(const MeshView& m, const NListView& cl, double rlist, int* counts, hipStream_t s);
void launch_verlet_fill(const MeshView& m, const NListView& cl, double rlist, const int* offsets, int* adj,
                        hipStream_t s);

// flips (tm_flips.hip)
struct FlipMeshView {
    double* x;
    int* he_to;
    int* he_next;
    int* he_prev;
    int* he_face;
    int* v_he;
    int* f_he;
    int nv, nh, nf;
};
void launch_flip_serial(const FlipMeshView& fm, const NListView& nl, const EParams& p, int nflips,
                        const int* proposals, const double* uniforms, int n_uniforms, uint64_t seed, int injected,
                        const Globals* g0, int* result /* [accepted, used, error] */, hipStream_t s);

// graph-coloured (claim-based) parallel sweep; state receives props[6], the
// final energy and the number of accepted flips
void launch_flip_parallel(const FlipMeshView& fm, const NListView& nl, const EParams& p, int nflips, int batch,
                          uint64_t seed, const Globals* g0, unsigned long long* owner, int* status, double* dprops,
                          double* state, hipStream_t s);

// device HMC step (tm_hmc.hip)
struct HmcState {
    double energy0, energy1, kinetic0, kinetic1, dh, uniform;
    int accepted, pad;
    unsigned long long n_accepted, n_steps;
};
enum { kHmcFirst = 1, kHmcKick = 2, kHmcDrift = 4, kHmcLast = 8 };
int hmc_blocks(int n3);
void launch_hmc_begin(int n3, const double* x, double* x0, double* p, const int* perm, int draw, uint64_t seed,
                      uint64_t step, double sigma, double* kpart, hipStream_t s);
void launch_hmc_leap(int n3, double* x, double* p, double* a, const double* g, double dt, double dt2, double mass,
                     int mode, const double* energy, HmcState* st, double* kpart, hipStream_t s);
void launch_hmc_accept(int n3, const double* kpart0, const double* kpart1, const double* x0, const double* x,
                       int harmonic, const double* energy, double mass, double temperature, double uniform,
                       uint64_t seed, uint64_t step, HmcState* st, hipStream_t s);
void launch_hmc_restore(int n3, double* x, const double* x0, const HmcState* st, hipStream_t s);

}    // namespace tmh
