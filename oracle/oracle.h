/* oracle.h -- C API of the FP64 CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is a from-scratch CPU restatement of
 * bio-phys/trimem's energy / gradient / neighbour-list / flip arithmetic
 * (src/mesh_util.h, src/mesh_properties.cpp, src/kernel.h, src/energy.cpp,
 * src/mesh_tether.h, src/mesh_repulsion.h, src/nlists/{cell_list,verlet_list,nlist}.h, src/flips.cpp,
 * src/flip_utils.cpp, src/util.cpp, src/mesh_py.cpp) and of the OpenMesh
 * half-edge behaviour those files rely on.  Only tests/, the smoke check and
 * the CPU-baseline leg of bench.py may load it; the product library never
 * links or calls it.
 *
 * All functions return 0 on success and a negative code on error unless the
 * return value is a count or a value; orc_last_error() gives the message.
 */
#ifndef TRIMEM_ORACLE_H
#define TRIMEM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Mirror of trimem's EnergyParams / BondParams / SurfaceRepulsionParams /
 * ContinuationParams (src/params.h:19-71), flattened. */
typedef struct {
    double kappa_b, kappa_a, kappa_v, kappa_c, kappa_t, kappa_r;
    double area_frac, volume_frac, curvature_frac;
    double bond_lc0, bond_lc1, bond_a0;
    int bond_r;
    int bond_type;        /* 0 Edge, 1 Area, 2 None (params.h:13-18) */
    double rep_lc1;
    int rep_r;
    int rep_nlist;        /* 0 "cell-list", 1 "verlet-list" */
    double rep_rlist;
    int rep_exclusion;
    int pad_;
    double cont_delta, cont_lambda;
} orc_params;

const char* orc_last_error(void);
int orc_num_threads(void);

/* ---- mesh (OpenMesh TriMesh_ArrayKernelT subset) ---- */
void* orc_mesh_new(const double* points, int n_points, const int* cells, int n_cells);
void  orc_mesh_free(void* mesh);
int   orc_mesh_nv(void* mesh);
int   orc_mesh_ne(void* mesh);
int   orc_mesh_nf(void* mesh);
void  orc_mesh_get_points(void* mesh, double* out);
void  orc_mesh_set_points(void* mesh, const double* in);
void  orc_mesh_fv_indices(void* mesh, int* out);
void  orc_mesh_topology(void* mesh, int* he_to, int* he_next, int* he_face, int* v_he, int* f_he);
int   orc_mesh_is_flip_ok(void* mesh, int edge);
void  orc_mesh_flip(void* mesh, int edge);
int   orc_mesh_next_halfedge(void* mesh, int he);

/* ---- geometry primitives on one halfedge (mesh_util.h:15-185) ---- */
double orc_edge_length(void* mesh, int he);
void   orc_edge_length_grad(void* mesh, int he, double* out /*2x3*/);
double orc_face_area(void* mesh, int he);
void   orc_face_area_grad(void* mesh, int he, double* out /*3x3*/);
double orc_face_volume(void* mesh, int he);
void   orc_face_volume_grad(void* mesh, int he, double* out /*3x3*/);
double orc_dihedral_angle(void* mesh, int he);
void   orc_dihedral_angle_grad(void* mesh, int he, double* out /*4x3*/);

/* ---- energy manager (energy.cpp:13-151) ---- */
void*  orc_energy_new(void* mesh, const orc_params* p);
void   orc_energy_free(void* em);
int    orc_properties(void* em, void* mesh, double* out6);
int    orc_vertex_properties(void* em, void* mesh, double* out /* n x 6 */);
double orc_energy(void* em, void* mesh);
double orc_energy_props(void* em, const double* props6);
int    orc_gradient(void* em, void* mesh, double* out /* n x 3 */);
void   orc_update_reference_properties(void* em);
int    orc_update_repulsion(void* em, void* mesh);
void   orc_initial_props(void* em, double* out6);
void   orc_get_params(void* em, orc_params* out);
int    orc_fd_gradient(void* em, void* mesh, double* out, double eps);

/* ---- neighbour lists (nlists/nlist.cpp:10-53) ---- */
void*  orc_nlist_new(void* mesh, const orc_params* p);
void   orc_nlist_free(void* nl);
/* returns the number of entries; fills at most cap entries */
long   orc_nlist_distance_matrix(void* nl, void* mesh, double dmax, double* d, int* i, int* j, long cap);
long   orc_nlist_point_distances(void* nl, void* mesh, int pid, double dmax, double* dist3, int* jdx, long cap);
/* the energy manager's own list */
void*  orc_energy_nlist(void* em);

/* ---- flips (flips.cpp:22-80) with injected proposal and uniform streams ---- */
int    orc_flip_serial_injected(void* em, void* mesh, int nflips, const int* proposals,
                                const double* uniforms, int n_uniforms, int* n_used);
int    orc_edge_vertex_properties(void* em, void* mesh, int edge, double* out6);

/* ---- utilities (util.cpp:9-41) ---- */
double orc_area(void* mesh);
double orc_edges_length(void* mesh);

#ifdef __cplusplus
}
#endif
#endif
