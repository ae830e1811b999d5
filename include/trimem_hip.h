/* trimem_hip.h -- C ABI of the MI355X (gfx950) trimem energy / gradient /
 * neighbour-list / flip library (libtrimem_hip.so).
 *
 * Drop-in boundary for bio-phys/trimem's compiled core: every entry point
 * below replaces one C++ entry of src/ that src/bindings.cpp exposes as
 * trimem.core (bindings.cpp:1038-1093); the citation after each declaration
 * names it.  Plain pointers and sizes only; no exception crosses the ABI.
 *
 * Conventions
 *   - Every function returns TM_OK (0) or a negative TM_E* code; the message
 *     of the last failure on the calling thread is tm_last_error().
 *   - Host arrays are row-major and caller-owned: points double[n*3],
 *     faces int[m*3], gradients double[n*3], exactly numpy's (n,3) C order.
 *   - A tm_mesh owns a host points buffer (tm_mesh_points) that the caller
 *     may write at any time, like trimem's TriMesh.points() view
 *     (mesh_py.cpp:96-105); every host-facing compute call uploads it first.
 *   - The *_device entry points work on the device-resident positions
 *     (tm_mesh_device_points) and leave results in device memory; they do
 *     not synchronise.
 *   - Handles are not thread-safe: one thread at a time per handle, which is
 *     how the reference is used under the Python GIL.
 */
#ifndef TRIMEM_HIP_H
#define TRIMEM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    TM_OK = 0,
    TM_E_INVALID = -1,   /* bad argument or shape (reference: std::runtime_error)  */
    TM_E_RANGE = -2,     /* flip_ratio > 1 for the parallel flips (std::range_error) */
    TM_E_DEVICE = -3,    /* HIP runtime failure or no device                        */
    TM_E_UNSUPPORTED = -4
};

typedef struct tm_mesh tm_mesh;
typedef struct tm_energy tm_energy;
typedef struct tm_nlist tm_nlist;

/* Flattened EnergyParams / BondParams / SurfaceRepulsionParams /
 * ContinuationParams (src/params.h:11-71). */
typedef struct {
    double kappa_b, kappa_a, kappa_v, kappa_c, kappa_t, kappa_r;
    double area_frac, volume_frac, curvature_frac;
    double bond_lc0, bond_lc1, bond_a0;
    int bond_r;
    int bond_type;      /* 0 Edge, 1 Area, 2 None (params.h:13-18)            */
    double rep_lc1;
    int rep_r;
    int rep_nlist;      /* 0 "cell-list", 1 "verlet-list" (nlist.cpp:14-50) */
    double rep_rlist;
    int rep_exclusion;  /* 0, 1 or 2                                         */
    int reserved;
    double cont_delta, cont_lambda;
} tm_params;

/* TVertexProperties<real> (mesh_properties.h:15-48) */
typedef struct {
    double area, volume, curvature, bending, tethering, repulsion;
} tm_props;

enum { TM_PREC_FP64 = 0, TM_PREC_MIXED = 1 };

const char* tm_last_error(void);
int tm_device_count(int* count);
int tm_set_device(int device);

/* ---- mesh: TriMesh (mesh.h:15-31, mesh_py.cpp:12-105, bindings.cpp:52-193) ---- */
/* TriMesh(points, faces) -> from_points_cells (mesh_py.cpp:12-69) */
int tm_mesh_create(const double* points, int n_points, const int* faces, int n_faces, tm_mesh** out);
int tm_mesh_destroy(tm_mesh* mesh);
/* n_vertices / n_edges / n_faces (bindings.cpp:136-146) */
int tm_mesh_counts(const tm_mesh* mesh, int* n_vertices, int* n_edges, int* n_faces);
/* points() view (mesh_py.cpp:96-105): host buffer owned by the mesh */
double* tm_mesh_points(tm_mesh* mesh);
/* fv_indices (mesh_py.cpp:73-93) */
int tm_mesh_fv_indices(tm_mesh* mesh, int* out);
/* halfedge_handle(i) / next_halfedge_handle / opposite_halfedge_handle
 * (bindings.cpp:148-174); opposite is h ^ 1 */
int tm_mesh_next_halfedge(tm_mesh* mesh, int he, int* out);
/* halfedge arrays (to, next, face per halfedge; halfedge per vertex/face) */
int tm_mesh_topology(tm_mesh* mesh, int* he_to, int* he_next, int* he_face, int* v_he, int* f_he);
/* device-resident positions: double[n*3] in HBM */
double* tm_mesh_device_points(tm_mesh* mesh);
/* host -> device (the host array wins) and device -> host copies of the
 * positions */
int tm_mesh_upload_points(tm_mesh* mesh);
int tm_mesh_download_points(tm_mesh* mesh);
/* After a device-resident step (tm_hmc_step) the device positions are ahead
 * of the host array: compute calls then use the device copy as it is, and
 * this call copies it back to the host array (no-op otherwise).  Call it
 * before reading or writing tm_mesh_points after device-resident steps. */
int tm_mesh_sync_host(tm_mesh* mesh);

/* ---- geometry on one halfedge (mesh_util.h:15-185, bindings.cpp:244-407) ----
 * out[47] = [edge_length, edge_length_grad(2x3), face_area, face_area_grad(3x3),
 *            face_volume, face_volume_grad(3x3), dihedral_angle,
 *            dihedral_angle_grad(4x3)] */
int tm_halfedge_geometry(tm_mesh* mesh, int he, double* out47);

/* ---- utilities (util.cpp:9-41, bindings.cpp:987-1036) ---- */
int tm_area(tm_mesh* mesh, double* out);
int tm_edges_length(tm_mesh* mesh, double* out);
/* No counterpart in the reference: tm_area (what = 0) or tm_edges_length
 * (what = 1) for callers that cannot allocate device memory -- code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.
 * It sums over the device-resident positions (tm_mesh_device_points; the
 * mesh must already be on the device, e.g. after tm_mesh_upload_points),
 * keeps its block partials and arrival counter in a fixed scratch array of
 * the library (one per GPU, held until the process ends; meshes of at most
 * 8388608 faces or edges, larger ones are refused with TM_E_INVALID), and
 * writes the sum to *out, a double in device memory.  It calls no hipMalloc
 * or hipFree and does not synchronise (tm_sync, or the library stream from
 * tm_stream, orders the result).  All calls share that scratch: calls that
 * overlap on one GPU (from several threads) race. */
int tm_mesh_sum_noalloc(tm_mesh* mesh, int what, double* out);

/* ---- energy: EnergyManager (energy.h:18-56, energy.cpp:13-151) ---- */
/* EnergyManager(mesh, eparams) (energy.cpp:13-28); bond type None and
 * unknown list types or exclusion levels fail with TM_E_INVALID
 * (mesh_tether.h:123-135, nlist.cpp:10-53) */
int tm_energy_create(tm_mesh* mesh, const tm_params* params, tm_energy** out);
int tm_energy_destroy(tm_energy* em);
int tm_energy_get_params(const tm_energy* em, tm_params* out);
int tm_energy_initial_props(const tm_energy* em, tm_props* out);
/* precision of properties/energy/gradient: TM_PREC_FP64 (default) or mixed */
int tm_energy_set_precision(tm_energy* em, int precision);
int tm_energy_get_precision(const tm_energy* em, int* precision);
/* properties(mesh) -> reduced VertexProperties (energy.cpp:63-77) */
int tm_energy_properties(tm_energy* em, tm_mesh* mesh, tm_props* out);
/* per-vertex VertexProperties vector (EvaluateProperties, kernel.h:140-169) */
int tm_energy_vertex_properties(tm_energy* em, tm_mesh* mesh, double* out /* n*6 */);
/* energy(mesh) (energy.cpp:79-85) */
int tm_energy_energy(tm_energy* em, tm_mesh* mesh, double* out);
/* energy(vprops) (energy.cpp:87-91) */
int tm_energy_energy_props(tm_energy* em, const tm_props* props, double* out);
/* gradient(mesh) (energy.cpp:93-124) */
int tm_energy_gradient(tm_energy* em, tm_mesh* mesh, double* out /* n*3 */);
/* energy(mesh) and gradient(mesh) from one properties pass */
int tm_energy_energy_gradient(tm_energy* em, tm_mesh* mesh, double* energy, double* grad /* n*3 */);
/* update_reference_properties (energy.cpp:47-55) */
int tm_energy_update_reference_properties(tm_energy* em);
/* update_repulsion (energy.cpp:57-61) */
int tm_energy_update_repulsion(tm_energy* em, tm_mesh* mesh);
/* device-resident evaluation on tm_mesh_device_points (no counterpart in the
 * reference, whose gradient() always starts from host points): flags bit 0
 * runs the properties pass (energy, globals), bit 1 the gradient pass (which
 * needs a properties pass before it).  Enqueued on the library stream
 * (tm_stream), no synchronisation; results via the pointers below. */
enum { TM_EVAL_PROPS = 1, TM_EVAL_GRAD = 2 };
int tm_energy_eval_device(tm_energy* em, tm_mesh* mesh, int flags);
const double* tm_energy_device_gradient(tm_energy* em);   /* double[n*3] */
const double* tm_energy_device_globals(tm_energy* em);    /* props[6], energy */
int tm_sync(void);
/* the library's HIP stream (hipStream_t), for event timing by callers */
int tm_stream(void** out);
/* timing aid (no counterpart in the reference; for callers that time one
 * kernel of a launch sequence with HIP events): the next properties pass
 * (tm_energy_eval_device or tm_replicas_eval with TM_EVAL_PROPS, or any
 * single-mesh call that runs one) stamps the events `before` and `after`
 * (hipEvent_t; either may be null) at the start and the end of its pass-1
 * kernel (one launch over the whole group for tm_replicas_eval;
 * hipExtLaunchKernel), then forgets them.
 * The pass does the same work with or without them. */
int tm_mark_pass1(void* before, void* after);

/* ---- neighbour lists (nlist.h:15-24, nlist.cpp:10-53, bindings.cpp:893-985) ---- */
/* make_nlist(mesh, eparams) */
int tm_nlist_create(tm_mesh* mesh, const tm_params* params, tm_nlist** out);
int tm_nlist_destroy(tm_nlist* nl);
/* the energy manager's own list (EnergyManager::nlist), borrowed: after
 * tm_energy_update_repulsion it is the previous list (the next rebuild
 * reuses its buffers), so ask again */
int tm_energy_nlist(tm_energy* em, tm_nlist** out);
/* distance_matrix(mesh, dmax): count first (d, i, j may be NULL), then fill */
int tm_nlist_distance_matrix(tm_nlist* nl, tm_mesh* mesh, double dmax, double* d, int* i, int* j, long cap,
                             long* count);
/* point_distances(mesh, pid, dmax) */
int tm_nlist_point_distances(tm_nlist* nl, tm_mesh* mesh, int pid, double dmax, double* dist3, int* jdx,
                             long cap, long* count);

/* ---- flips (flips.h:11-17, bindings.cpp:839-891) ---- */
/* flip(mesh, estore, flip_ratio) with the random stream drawn from seed
 * (the reference draws from a clock-seeded mt19937, flips.cpp:20) */
int tm_flip_serial(tm_energy* em, tm_mesh* mesh, double flip_ratio, uint64_t seed, int* accepted);
/* pflip(mesh, estore, flip_ratio) (flip_parallel_batches, flips.cpp:82-166):
 * rounds of `batch` proposals (<= 0: automatic); in each round proposals
 * claim their flip patch (flip_utils.cpp:43-68), the winners (disjoint
 * patches) flip in parallel and a sequential FP64 Metropolis scan in
 * proposal order accepts or rejects them.  flip_ratio > 1 -> TM_E_RANGE. */
int tm_flip_parallel(tm_energy* em, tm_mesh* mesh, double flip_ratio, uint64_t seed, int batch, int* accepted);
/* the last parallel sweep's own bookkeeping (tm_flip_parallel or
 * tm_replicas_pflip on this energy manager; no counterpart in the
 * reference): the running global properties (area, volume, curvature,
 * bending, tethering, repulsion), the energy and the accepted count, as the
 * Metropolis scan left them.  The scan updates them by each accepted flip's
 * patch difference, so they equal the properties of the final topology when
 * no two flips accepted in one round shared a patch edge.  A sweep with no
 * proposals leaves them as they were; all zero before the first sweep. */
int tm_energy_pflip_state(tm_energy* em, double* out /* 8 */);
/* flip_serial with caller-supplied proposals (edge ids) and uniforms; a
 * uniform is consumed only by proposals that pass is_flip_ok */
int tm_flip_serial_injected(tm_energy* em, tm_mesh* mesh, int n_proposals, const int* proposals,
                            const double* uniforms, int n_uniforms, int* accepted, int* n_used);

/* ---- device-resident HMC step (trimem/mc/hmc.py:14-26, 129-158) ----
 * One step of trimem's HMC sampler with positions, momenta and forces in HBM:
 * momenta p ~ N(0, mass*T), num_integration_steps velocity-Verlet steps
 * (hmc.py:14-26, forces from the energy passes), FP64 Metropolis test on
 * H = E + 0.5 p.p / mass (hmc.py:129-131, 146-155), positions restored on
 * rejection.  Enqueued on the library stream; no host synchronisation. */
typedef struct tm_hmc tm_hmc;
enum { TM_HMC_ENERGY = 0,      /* E = the energy manager's energy */
       TM_HMC_HARMONIC = 1 };  /* E = 0.5 x.x (tests/test_sampling.py:40-44) */
typedef struct tm_hmc_params {
    double mass;               /* "mass" (hmc.py:38) */
    double time_step;          /* "time_step" (hmc.py:39) */
    double time_step_sq;       /* time_step squared as the caller forms it
                                  (hmc.py:21 uses dt**2); 0 -> dt*dt */
    int num_integration_steps; /* L >= 0 (hmc.py:40) */
    int reserved;
} tm_hmc_params;
typedef struct tm_hmc_outcome {
    double dh;                 /* (H(xn,pn) - H(x,p)) / T of the last step */
    double energy0, energy1;   /* E(x), E(xn) */
    double kinetic0, kinetic1; /* 0.5 p.p / mass at start and end */
    double uniform;            /* the uniform compared with min(1, exp(-dh)) */
    long long accepted;        /* last step: 1 accepted, 0 rejected */
    long long n_accepted;      /* accepted steps since creation */
    long long n_steps;
} tm_hmc_outcome;
/* em may be NULL for TM_HMC_HARMONIC; mesh and em must outlive the handle */
int tm_hmc_create(tm_mesh* mesh, tm_energy* em, int model, tm_hmc** out);
int tm_hmc_destroy(tm_hmc* hmc);
/* momenta: double[n*3] in user vertex order, the caller's
 * normal(size)*sqrt(mass*T) draw (host RNG, hmc.py:140), or NULL for a
 * counter-based device draw keyed by (seed, step).  uniform: the caller's
 * U[0,1) draw (hmc.py:151), or < 0 for a device draw keyed by (seed, step).
 * Host momenta are copied before the call returns. */
int tm_hmc_step(tm_hmc* hmc, const tm_hmc_params* params, double temperature, const double* momenta,
                double uniform, uint64_t seed, uint64_t step);
/* result of the last step and running counts (synchronises) */
int tm_hmc_result(tm_hmc* hmc, tm_hmc_outcome* out);

/* ---- launch shapes and the autotuner ----
 * No counterpart in the reference (OpenMP schedules).  The repulsion query
 * and the gradient pass write per-vertex values only, so their block size and
 * the number of candidates each lane keeps in flight change no result bit.
 * The query may also share one launch with the properties pass (rep_fused):
 * its blocks then sum the repulsion column of the block partials in the
 * order the properties pass uses, so this changes no bit either; it pays on
 * small meshes, whose launches do not fill the GPU.  The autotuner times
 * every shape on the caller's mesh (HIP events, the mean of 5 launches after
 * 2 untimed ones), then the fused form against the best separate one, and
 * keeps the fastest.  Replica groups launch the query apart (their
 * rep_fused reads 0); their pair cache keeps a recording flag and a
 * displacement check per replica on the device, so graph replays stay valid. */
typedef struct tm_tuning {
    int rep_threads;     /* repulsion query: 64, 128, 256 or 512 threads per block */
    int rep_batch;       /* candidates in flight per lane: 1, 2, 4 or 8 */
    int pass2_threads;   /* gradient pass: 64, 128, 256 or 512 */
    int rep_fused;       /* 1: the query shares the launch of the properties pass
                            (256 threads, rep_threads unused), 0: its own launch */
    int pair_cache;      /* 1 (default): the cell-list query replays recorded
                            candidate lists while they are provably complete
                            (single meshes; bit-identical), 0: full query */
    int pflip_window;    /* 1..256 (default 256): proposals the parallel flip
                            sweep tests at once against the running state when
                            it looks for the next acceptance; every window
                            takes the decisions of the sequential walk (1) */
    int serial_window;   /* 1..256 (default 256): proposals the serial flip sweep
                            evaluates at once before committing them in order;
                            0: one workgroup walks the proposals.  Every value
                            gives the same accepted flips and topology */
} tm_tuning;
int tm_energy_set_tuning(tm_energy* em, const tm_tuning* t);
int tm_energy_get_tuning(const tm_energy* em, tm_tuning* out);
/* ms_rep / ms_pass2 (may be NULL): the winners' times per launch (ms_rep:
 * of the fused properties launch when rep_fused is chosen) */
int tm_energy_autotune(tm_energy* em, tm_mesh* mesh, tm_tuning* out, double* ms_rep, double* ms_pass2);

/* ---- replica groups: batched meshes, parallel tempering, graph replay ----
 * No counterpart in the reference, which evaluates one mesh per call
 * (kernel.h:140-262).  R meshes with their energy managers, of any sizes,
 * advance together: every kernel of an evaluation or of an HMC step is one
 * launch over all replicas, and each replica's results equal, bit for bit,
 * those of the single-mesh calls (tm_energy_eval_device, tm_hmc_step with
 * device draws).  All energy managers must use the same precision.  Like
 * tm_energy_eval_device, the group works on the device positions: creation
 * uploads the host arrays that are newer, later host writes need
 * tm_mesh_upload_points. */
typedef struct tm_replicas tm_replicas;
int tm_replicas_create(tm_mesh* const* meshes, tm_energy* const* ems, int n_replicas, tm_replicas** out);
int tm_replicas_destroy(tm_replicas* rs);
/* TM_EVAL_PROPS | TM_EVAL_GRAD for every replica; results in each energy
 * manager's device buffers (tm_energy_device_globals / _gradient) */
int tm_replicas_eval(tm_replicas* rs, int flags);
/* pflip (flip_parallel_batches, flips.cpp:82-166) on every replica at once:
 * replica r draws its proposals from seeds[r] and flips with the kappa_b of
 * the ladder slot it holds; each replica's flips and accepted count are
 * those of tm_flip_parallel(ems[r], meshes[r], flip_ratio, seeds[r], 0, .)
 * (round k of every replica runs in one launch per phase).  accepted:
 * int[n_replicas] or NULL.  flip_ratio > 1 -> TM_E_RANGE. */
int tm_replicas_pflip(tm_replicas* rs, double flip_ratio, const uint64_t* seeds, int* accepted);
/* out[7 r + k]: the six global properties (area, volume, curvature,
 * bending, tethering, repulsion) and the energy of the last evaluation of
 * replica r (synchronises) */
int tm_replicas_globals(tm_replicas* rs, double* out);
/* gradient of the last evaluation of replica r, double[n_r * 3] in user
 * vertex order (synchronises) */
int tm_replicas_gradient(tm_replicas* rs, int r, double* out);
/* HMC parameters per replica (arrays of n_replicas; seeds key the device
 * draws, see tm_hmc_step) and the shared number of leapfrog steps */
int tm_replicas_hmc_setup(tm_replicas* rs, const double* mass, const double* time_step, const double* temperature,
                          const uint64_t* seeds, int num_integration_steps);
/* kappa_b ladder (n_replicas values, ascending) for replica exchange: slot
 * s holds ladder[s]; replica r starts on slot r and is evaluated with the
 * kappa_b of its current slot */
int tm_replicas_set_ladder(tm_replicas* rs, const double* ladder, uint64_t seed);
/* n_steps HMC steps of every replica, momenta and uniforms drawn on the
 * device from (seed_r, step count of replica r); swap_every > 0: one
 * exchange round (alternating even and odd slot pairs) after every
 * swap_every-th step; use_graph != 0: each step replayed from one captured
 * hipGraph.  Enqueued on the library stream, no host synchronisation. */
int tm_replicas_hmc_run(tm_replicas* rs, int n_steps, int swap_every, int use_graph);
/* one exchange round now */
int tm_replicas_swap(tm_replicas* rs);
/* per replica: accepted steps, steps, current ladder slot, and the six
 * global properties of its current state (props: 6 per replica); per slot
 * s (pair s, s+1): exchanges accepted and tried.  Any pointer may be NULL.
 * Synchronises. */
int tm_replicas_stats(tm_replicas* rs, long long* accepted, long long* steps, int* slot_of, double* props,
                      long long* swaps_accepted, long long* swaps_tried);
/* rebuild every replica's neighbour list from its current device positions
 * (EnergyManager.update_repulsion, energy.cpp:54-61) */
int tm_replicas_update_repulsion(tm_replicas* rs);
/* launch shapes of the grouped launches (see tm_energy_autotune) */
int tm_replicas_set_tuning(tm_replicas* rs, const tm_tuning* t);
int tm_replicas_get_tuning(const tm_replicas* rs, tm_tuning* out);
int tm_replicas_autotune(tm_replicas* rs, tm_tuning* out, double* ms_rep, double* ms_pass2);

/* ---- surrogate forces (no counterpart in the reference) ----
 * A one-ring message-passing network (width 64, three layers, bf16 MFMA
 * products with FP32 accumulation) that predicts F = -dE/dx from invariant
 * ring features; trained on the GPU by its own forward and backward kernels
 * from (positions, exact gradient) snapshots.  HMC steps driven by it are
 * accepted with the exact energy, so sampling stays exact.  Closed meshes
 * with valences 3..10. */
typedef struct tm_surrogate tm_surrogate;
/* random initial weights from seed; feature normalisation and force scale
 * from the mesh's current state */
int tm_surrogate_create(tm_mesh* mesh, tm_energy* em, uint64_t seed, tm_surrogate** out);
int tm_surrogate_destroy(tm_surrogate* sg);
/* the FP32 parameter vector (out may be NULL to query n) */
int tm_surrogate_params(tm_surrogate* sg, float* out, int* n);
int tm_surrogate_set_params(tm_surrogate* sg, const float* in);
/* feature means and scales (8 each), mean edge length, force scale */
int tm_surrogate_norm(tm_surrogate* sg, double* mu, double* sig, double* lbar, double* sigma_F);
/* predicted forces at the current positions, double[n*3] user order (may be NULL) */
int tm_surrogate_forces(tm_surrogate* sg, double* forces);
/* loss against the exact gradient at the current positions and its
 * gradient with respect to the parameters (grad may be NULL) */
int tm_surrogate_loss_grad(tm_surrogate* sg, double* loss, float* grad);
/* activations of the last forward pass, user vertex order: which 0 the
 * input features (n x 16), 1..4 the hidden states H_0..H_3, 5..7 the
 * aggregated messages M_1..M_3 (n x 64) */
int tm_surrogate_activations(tm_surrogate* sg, int which, float* out);
/* record the current positions and their exact gradient on the device */
int tm_surrogate_add_snapshot(tm_surrogate* sg);
int tm_surrogate_num_snapshots(tm_surrogate* sg, int* n);
/* Adam (lr) over the snapshots, reshuffled (seed) every epoch; the mesh
 * positions are restored afterwards; mean_loss: mean loss of the last epoch */
int tm_surrogate_train(tm_surrogate* sg, int epochs, float lr, uint64_t seed, double* mean_loss);
/* tm_hmc_step with the surrogate's forces and the exact energies */
int tm_hmc_step_surrogate(tm_hmc* hmc, tm_surrogate* sg, const tm_hmc_params* params, double temperature,
                          const double* momenta, double uniform, uint64_t seed, uint64_t step);
/* One HMC step (trimem/mc/hmc.py:133-158) driven by the surrogate while it
 * earns its keep, by the exact gradient otherwise; every step is accepted
 * with the exact energy either way.  The acceptance of the last `window`
 * surrogate-driven steps is kept; when it falls below `min_acceptance` the
 * exact gradient (tm_hmc_step) drives, except one surrogate-driven probe
 * every `probe_every` steps, and once `window` probes accept at that rate
 * again the surrogate drives again.  Topologies the network cannot take
 * (boundary, valence above 10) are driven by the exact gradient.  driver
 * (may be NULL): 1 surrogate, 0 exact.  Reading a surrogate-driven step's
 * outcome makes the next guarded step wait for it.  Defaults: window 10,
 * min_acceptance 0.5, probe_every 10. */
int tm_hmc_step_guarded(tm_hmc* hmc, tm_surrogate* sg, const tm_hmc_params* params, double temperature,
                        const double* momenta, double uniform, uint64_t seed, uint64_t step, int* driver);
int tm_surrogate_set_policy(tm_surrogate* sg, int window, double min_acceptance, int probe_every);
/* falling_back: 1 while the exact gradient drives; counts of surrogate- and
 * exact-driven guarded steps and of switches; the acceptance over the
 * current window (-1 if empty).  Any pointer may be NULL. */
int tm_surrogate_policy_state(tm_surrogate* sg, int* falling_back, long long* n_surrogate, long long* n_exact,
                              long long* n_switches, double* recent_acceptance);

#ifdef __cplusplus
}
#endif
#endif
