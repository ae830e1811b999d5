// tm_internal.h -- launchers shared between the library's translation units.
#pragma once

#include "tm_common.h"

namespace tmh {

// Pair cache of the cell-list repulsion query (single meshes).  Between
// rebuilds of the cell list the candidates of a vertex are fixed (the stored
// stencil list of its cell, or the stencil rows of the cell it has moved
// to), and exclusion depends on the topology only.  A recording launch runs
// the full query with radius lc1 + skin and keeps, per vertex and lane, the
// non-excluded candidates it meets, in the lane's visit order; later
// launches replay those lists with the exact lc1 test while every vertex has
// moved less than skin / 2 since the recording (no other candidate can have
// come within lc1) and the vertex is in the cell it was recorded in (the
// full query would visit the same candidates in the same order).  Each lane
// then adds the same terms in the same order as the full query: the results
// are bit-identical.
constexpr int kPairLanes = 8;      // lanes per vertex of the query
constexpr int kPairCap = 4;        // recorded candidates per lane (more: that vertex takes the full query)
constexpr double kPairSkin = 0.2;  // skin = 0.2 lc1
// Replay while every vertex has moved less than 0.25 skin from the snapshot
// of the last recording of all lists.  A vertex that left the cell its list
// was recorded in records it again there (the snapshot stays), so two lists
// can be recorded at different times: a pair's distance then changed by at
// most 4 x 0.25 skin since either list was recorded.
constexpr double kPairLimit = 0.25;
struct PairCache {
    double* x_snap;         // [3N] positions at the recording; null: no cache
    int* cell;              // [N] cell id at the recording, -1: replay not allowed
    unsigned* cnt;          // [N] recorded candidates of each lane, 4 bits per lane
    int* ent;               // [8 kPairCap N] ids, lane-major per vertex
    double* disp_part;      // [disp_n] max |x - x_snap|^2 per pass-1 block (written by pass 1)
    int disp_n;
    int* dforce;            // replica groups: device flag, nonzero records (the reduction clears it)
    double skin;
    double limit2;          // replay while max |x - x_snap|^2 < limit2
    int force;              // single meshes: 1 records in this launch
};
constexpr int kCachedVerts = 256;    // vertices per block of the cached query
constexpr int kPass1Verts = 256;     // vertices (one lane each) per block of pass 1

// energy passes (tm_energy.hip)
struct PassBuffers {
    double* q;          // [N] curvature / area per vertex
    double* gprops;     // [6N] properties of listed (boundary) vertices
    int* pair_cnt;      // [N] repulsion pairs per vertex (k_rep)
    double* rep_v;      // [N] repulsion property per vertex (k_rep)
    double* rep_g;      // [3N] repulsion gradient, rows of vertices with pairs (k_rep)
    double* partials;   // [6 * pass1 blocks]
    Globals* globals;
    const int* list;    // boundary / degenerate vertices
    int nlist;
    PairCache pc;       // x_snap null: the query runs in full every time
    // [kRingPad * N] dihedral angle of each ring slot of vertices of valence
    // <= kRingPad, column-major, in the precision of the pass (float or
    // double); written by pass 1, read by pass 2.  Null: pass 2 forms them
    void* theta;
};
int pass1_blocks(int nv);

// Everything the pass kernels read and write for one mesh.  A grouped launch
// takes a device array of these, one per replica, and a prefix table of
// pass-1 blocks: logical block b belongs to the replica r with
// block0[r] <= b < block0[r + 1] and is its local block b - block0[r], so
// every replica is partitioned and summed exactly as in a launch of its own.
struct PassArgs {
    MeshView m;
    NListView nl;
    EParams p;
    double* q;
    double* gprops;
    double* rep_v;       // null: repulsion off
    int* pair_cnt;
    double* rep_g;       // null: rows not stored (pass 1 only)
    double* partials;
    Globals* globals;
    double* grad;
    double* vprops;      // optional per-vertex properties of pass 1
    void* theta;         // PassBuffers::theta
    const int* list;     // boundary / degenerate vertices
    int nlist;
    int nblocks;         // pass-1 blocks of kPass1Verts vertices (a multiple of 8)
    PairCache pc;
};
PassArgs pass_args(const MeshView& m, const NListView& nl, const EParams& p, const PassBuffers& b, double* vprops,
                   double* grad);
// Launch shapes that do not change any result bit: the repulsion query and
// pass 2 write per-vertex values only, and each lane visits its candidates in
// the same order whatever the batch.  (Pass 1's blocks fix the order of the
// global sums and stay at 256 threads; fused, the query's blocks cover the
// same 32 vertices as pass-1 blocks and sum their column the same way.)
struct Tuning {
    int rep_threads = 256;    // 64, 128, 256 or 512
    int rep_batch = 2;        // candidates in flight per lane: 1, 2, 4 or 8
    int p2_threads = 256;     // 64, 128, 256 or 512
    int rep_fused = 0;        // 1: the query shares the pass-1 launch (256 threads), 0: its own launch
    int pair_cache = 1;       // 1: replay recorded candidate lists (single meshes, cell list)
    int pflip_window = 256;   // proposals the pflip Metropolis scan tests at once (1..256)
    int serial_window = 256;  // proposals per window of the serial sweep (1..256), 0: one workgroup
};
bool tuning_valid(const Tuning& t);
int tuned_blocks(int nv, int threads);    // blocks of 8-lane groups (the query), a multiple of 8
int pass2_blocks(int nv, int threads);    // blocks of one lane per vertex (pass 2), a multiple of 8
// the repulsion query and the pass-2 kernel alone (for timing)
void launch_rep(const PassArgs& a, const Tuning& t, hipStream_t s);
// the repulsion query and the pass-1 kernel (fused or apart), no finalize;
// marks (optional): HIP events stamped at the pass-1 kernel's start and end
// (either may be null)
void launch_pass1_body(const PassArgs& a, int mixed, const Tuning& t, hipStream_t s,
                       const hipEvent_t* marks = nullptr);
void launch_pass2_kernel(const PassArgs& a, int mixed, const Tuning& t, hipStream_t s);
// R replicas: d_args in device memory, h_args its host copy (for the
// launches of the boundary kernels), one prefix table of blocks per kernel
// shape
struct GroupLaunch {
    const PassArgs* d_args;
    const PassArgs* h_args;
    int R;
    const int* d_block0_p1;
    int total_p1;
    const int* d_block0_rep;
    int total_rep;
    const int* d_block0_p2;
    int total_p2;
    Tuning tune;
    int mixed;
    // with the pair cache (h_args[r].pc.x_snap set): the cached query's
    // prefix table (pass-1 blocks of each replica)
    const int* d_block0_rc;
    int total_rc;
};
void launch_pass1_grouped(const GroupLaunch& g, hipStream_t s, const hipEvent_t* marks = nullptr);
void launch_pass2_grouped(const GroupLaunch& g, hipStream_t s);
void launch_rep_grouped(const GroupLaunch& g, hipStream_t s);
void launch_pass1(const MeshView& m, const NListView& nl, const EParams& p, int mixed, const PassBuffers& b,
                  double* vprops, hipStream_t s, const Tuning& t = Tuning(), const hipEvent_t* marks = nullptr);
void launch_pass2(const MeshView& m, const NListView& nl, const EParams& p, int mixed, const PassBuffers& b,
                  double* grad, hipStream_t s, const Tuning& t = Tuning());
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
void launch_fill_ints(int* a, int na, int va, int* b, int nb, int vb, int* c, int nc, int vc, hipStream_t s);
void launch_box(const double* x, int n, double* part, hipStream_t s);
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
// exclusion masks of the stored stencil lists (tm_energy.hip); nl needs
// vcand and cand, its excl_mask is ignored
void launch_excl_mask(const MeshView& m, const NListView& nl, unsigned long long* mask, hipStream_t s);
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

// windowed serial sweep: the serial sweep's decisions and topology from
// windows of up to kSerialWindowMax proposals evaluated in parallel
constexpr int kSerialWindowMax = 256;
struct SerialState {
    int pos, done, used, acc, err, window;
    double props[6];
    double e0;
};
struct SerialWindow {
    int* e;        // [W] proposed edge
    int* valid;    // [W] 1 flippable, 0 not, -1 index out of range
    int* pv;       // [4W] vertices whose one-rings the evaluation read (the patch)
    int* fc;       // [2W] the edge's two faces
    double* dO;    // [6W] patch properties before the flip
    double* dN;    // [6W] after the flip
    int* nbn;      // [4W] one-ring sizes of the patch vertices (-1: above kMaxValence)
    int* nb;       // [4 kMaxValence W] their one-rings (exclusion level 2 only)
};
void launch_serial_init(const Globals* g0, const EParams& p, int nflips, SerialState* st, hipStream_t s);
void launch_serial_windows(const FlipMeshView& fm, const NListView& nl, const EParams& p, int nflips,
                           const int* proposals, const double* uniforms, int n_uniforms, uint64_t seed, int injected,
                           int W, SerialState* st, const SerialWindow& w, unsigned long long* vstamp,
                           unsigned long long* estamp, unsigned long long* fstamp, int n_windows, hipStream_t s);

// graph-coloured (claim-based) parallel sweep; state receives props[6], the
// final energy and the number of accepted flips
void launch_flip_parallel(const FlipMeshView& fm, const NListView& nl, const EParams& p, int nflips, int batch,
                          int window, uint64_t seed, const Globals* g0, unsigned long long* owner, int* status,
                          double* dprops, double* state, hipStream_t s);

// Proposals per 64-lane block of the parallel sweep's claim phase (two lanes
// per patch vertex).
constexpr int kClaimPerBlock = 8;

// The parallel sweep of many meshes at once: round r of every replica in one
// launch per phase.  Each replica's kernels run the single-mesh sweep's
// device functions on its own arguments, so its flips, decisions and state
// are those of launch_flip_parallel with the same arguments.
struct PflipArgs {
    FlipMeshView fm;
    NListView nl;
    EParams p;
    uint64_t seed;
    int nflips, batch;
    const Globals* g0;
    unsigned long long* owner;
    int* status;
    double* dold;    // [6 batch]
    double* dnew;    // [6 batch]
    double* state;   // [8]
};
struct PflipGroup {
    const PflipArgs* d_args;
    int R;
    int rounds;            // the most rounds of any replica
    int window;
    // prefix tables of blocks per replica: claims (16 proposals per block),
    // evaluations (one proposal per block), applications (128 per block)
    const int* d_claim0;
    int total_claim;
    const int* d_eval0;
    int total_eval;
    const int* d_apply0;
    int total_apply;
};
void launch_flip_parallel_grouped(const PflipGroup& g, hipStream_t s);

// replica of logical block lb in a grouped launch (block0 ascending,
// block0[0] = 0)
__device__ __forceinline__ int find_replica(const int* __restrict__ block0, int R, int lb)
{
    int lo = 0, hi = R - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (block0[mid] <= lb) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// device HMC step (tm_hmc.hip)
struct HmcState {
    double energy0, energy1, kinetic0, kinetic1, dh, uniform;
    int accepted, pad;
    unsigned long long n_accepted, n_steps;
    double props0[6];    // global properties at x0 (first force evaluation)
    double props[6];     // ... of the state after the step (x0 or xn)
};
enum { kHmcFirst = 1, kHmcKick = 2, kHmcDrift = 4, kHmcLast = 8 };
int hmc_blocks(int n3);
void launch_hmc_begin(int n3, const double* x, double* x0, double* p, const int* perm, int draw, uint64_t seed,
                      uint64_t step, double sigma, double* kpart, hipStream_t s);
void launch_hmc_leap(int n3, double* x, double* p, double* a, const double* g, double dt, double dt2, double mass,
                     int mode, const Globals* globals, HmcState* st, double* kpart, hipStream_t s);
void launch_hmc_accept(int n3, const double* kpart0, const double* kpart1, const double* x0, const double* x,
                       int harmonic, const Globals* globals, double mass, double temperature, double uniform,
                       uint64_t seed, uint64_t step, HmcState* st, hipStream_t s);
void launch_hmc_restore(int n3, double* x, const double* x0, const HmcState* st, hipStream_t s);

// One replica of a grouped HMC step: device draws keyed by (seed, the
// replica's own step count), forces from grad, energies from globals.
struct HmcArgs {
    int n3;
    int nblocks;          // hmc_blocks(n3)
    double* x;
    double* x0;
    double* p;
    double* a;
    const double* grad;
    Globals* globals;
    const int* perm;
    HmcState* st;
    double* kpart0;
    double* kpart1;
    double mass, dt, dt2, temperature;
    uint64_t seed;
};
void launch_hmc_begin_grouped(const HmcArgs* d, const int* d_block0, int R, int total_blocks, hipStream_t s);
void launch_hmc_leap_grouped(const HmcArgs* d, const int* d_block0, int R, int total_blocks, int mode,
                             hipStream_t s);
void launch_hmc_accept_grouped(const HmcArgs* d, int R, hipStream_t s);
void launch_hmc_restore_grouped(const HmcArgs* d, const int* d_block0, int R, int total_blocks, hipStream_t s);
// Parallel-tempering exchange over a kappa_b ladder: ladder slots s and s+1
// (s = parity, parity + 2, ...) swap their replicas with probability
// min(1, exp(-(k[s+1] - k[s]) (B_a - B_b) / T)), B the bending of the current
// states; pass parameters and the slot owners are updated in place.
void launch_pt_swap(PassArgs* d_args, const HmcArgs* d_hmc, int* owner, const double* ladder, int R,
                    uint64_t seed, unsigned long long* round, unsigned long long* swaps_accepted,
                    unsigned long long* swaps_tried, hipStream_t s);

}    // namespace tmh
