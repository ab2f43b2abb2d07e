/*
 * shaw_amd.h -- C ABI of the MI355X-native SHAW time-stepping library.
 *
 * The reference (Pressio/SHAW) runs its hot path as C++ template calls inside
 * one executable; these entry points replace those calls one for one:
 *
 *   shaw_app_*        ShWavePP operator assembly + mesh files
 *                     (src/kokkos/shwavepp.hpp:100-148, 246-466;
 *                      src/shared/mesh_helpers/)
 *   shaw_map_*        source / receiver mapping
 *                     (src/shared/nominal_to_grid_mappers/)
 *   shaw_fom_*        runFom + updateVelocity/updateStress + forcing +
 *                     seismogram/snapshot collection
 *                     (src/kokkos/fom_run.hpp:57-161,
 *                      src/kokkos/fom_update_kernels.hpp:60-181,
 *                      src/kokkos/forcing_rank_{one,two}.hpp,
 *                      src/kokkos/seismogram.hpp:192-214,
 *                      src/kokkos/state_observer.hpp:174-203)
 *   shaw_rom_*        Galerkin projection and reduced leapfrog
 *                     (src/kokkos/rom_compute_jacobians.hpp:58-109,
 *                      src/kokkos/rom_run_rank_one.hpp:83-162,
 *                      src/kokkos/torevise/run_rom_rank_two_forcing.hpp:149-188,
 *                      src/tools/main_reconstruct_seismo.cc:13-41)
 *
 * Conventions: every function returns SHAW_OK (0) or an error code and the
 * message is available from shaw_last_error() (thread-local). No C++
 * exception crosses this boundary. Handles own their device buffers; host
 * buffers passed in are caller-owned and only read/written during the call.
 * Each FOM/ROM handle owns one HIP stream; calls on one handle are not
 * re-entrant. Arrays are C order unless stated; "LayoutLeft [a x b x c]"
 * means the reference's Kokkos::LayoutLeft order (first index fastest),
 * which equals C order [c][b][a].
 */
#ifndef SHAW_AMD_H_
#define SHAW_AMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SHAW_OK 0
#define SHAW_ERR_ARG 1     /* invalid argument or input file content */
#define SHAW_ERR_HIP 2     /* HIP runtime failure (no device, launch error) */
#define SHAW_ERR_MEMORY 3  /* allocation failed or footprint over budget */
#define SHAW_ERR_STATE 4   /* call out of order (e.g. advance before forcing) */

/* Message of the last failed call on this thread ("" if none). */
const char* shaw_last_error(void);
/* Library version string. */
const char* shaw_version(void);

/* ------------------------------------------------------------------ app --
 * Mesh + material -> matrix-free operators (the ShWavePP replacement).   */
typedef struct shaw_app* shaw_app_t;

/* Material kinds and their `params` flattening:
 *   0 unilayer : rho c0,c1,c2, vs c0,c1,c2  (polynomials in depth [m])
 *   1 bilayer  : depth2 [km], layer1 rho c0..c2, vs c0..c2, layer2 rho c0..c2, vs c0..c2
 *   2 prem     : (none)
 *   3 layered  : n, radius_0..radius_{n-1} [km, ascending], rho_0.. [kg/m^3], vs_0.. [m/s]
 *                (piecewise linear in radius between nodes, constant outside)
 * material_model_unilayer.hpp:72-83, material_model_bilayer.hpp:74-101,
 * material_model_prem.hpp:85-176.                                          */
typedef struct {
  int kind;
  int nparams;
  const double* params;
} shaw_material;

typedef struct {
  uint64_t nr, nth, nvp, nsp;
  double th_left_deg, th_right_deg, r_cmb_m, r_surf_m;
  double dth, drr;               /* [rad], [m] */
  double min_vs, max_vs;         /* over stress points, as ShWavePP */
} shaw_app_info;

/* Read mesh_info.dat / graph_vp.dat / coeff_vp.dat / graph_sp.dat from
 * mesh_dir (ShWavePP ctor, shwavepp.hpp:100-148). Fails if the graph is not
 * the structured full mesh meshing/ writes.                                */
int shaw_app_create_from_mesh(const char* mesh_dir, const shaw_material* mat, shaw_app_t* out);
/* Same operators for a mesh generated in memory (angles deg, radii km). */
int shaw_app_create_structured(uint64_t nr, uint64_t nth, double th_left_deg, double th_right_deg,
                               double r_cmb_km, double r_surf_km, const shaw_material* mat,
                               shaw_app_t* out);
/* Write the four mesh files exactly as meshing/create_single_mesh.py does. */
int shaw_mesh_write(const char* dir, uint64_t nr, uint64_t nth, double th_left_deg,
                    double th_right_deg, double r_cmb_km, double r_surf_km);
int shaw_app_info_get(shaw_app_t app, shaw_app_info* info);
/* coef_vp [nvp][4] (slots N,S,W,E), coef_sp [nsp][2], rho_inv [nvp], shear [nsp];
 * any pointer may be NULL.                                                */
int shaw_app_operators(shaw_app_t app, double* coef_vp, double* coef_sp, double* rho_inv,
                       double* shear_mod);
/* coords [n][2] = (theta [rad], r [m]) for dof 0 = vp, 1 = sp. */
int shaw_app_coords(shaw_app_t app, int dof, double* coords);
/* Nearest vp gid (ties -> last gid); sources on the theta edges move one
 * cell inward (map_point_source_to_velocity_grid_point.hpp:61-133,
 * map_nominal_location_to_velocity_grid_point.hpp:57-117).                */
int shaw_map_source(shaw_app_t app, double depth_km, double angle_deg, uint64_t* vp_gid);
int shaw_map_receiver(shaw_app_t app, double angle_deg, uint64_t* vp_gid);
/* FOM energy weights (no counterpart in the reference, whose ROM is
 * Euclidean): mv [nvp], ms [nsp] > 0 with Mv J_vp = -(Ms J_sp)^T on every
 * unmerged coupling, normalised to 1 at the mesh centre; the leapfrog
 * conserves 1/2 v^T Mv v + 1/2 s^T Ms s. For shaw_fom_pod_weighted and
 * shaw_rom_project_energy. Either pointer may be NULL.                     */
int shaw_app_energy_weights(shaw_app_t app, double* mv, double* ms);
void shaw_app_destroy(shaw_app_t app);

/* Signal kinds (signal.hpp:86-128): 0 ricker, 1 sinusoid, 2 gaussDer.
 * table[(s-1)*m + k] = f_k((s-1) dt), s = 1..nsteps, on the host with libm. */
int shaw_forcing_table(int m, const int* kinds, const double* delays, const double* periods,
                       double dt, uint64_t nsteps, double* table);

/* ---------------------------------------------------- device-side setup --
 * The setup the reference does on the host (material evaluation and
 * operator assembly, shwavepp.hpp:246-466; forcing values, signal.hpp:86-128;
 * source/receiver mapping, map_point_source_to_velocity_grid_point.hpp:61-133
 * and map_nominal_location_to_velocity_grid_point.hpp:57-117) run on the GPU,
 * one thread per point, step or location, from the mesh's per-row and
 * per-column tables (checked against every point of the mesh once; a mesh
 * that is not separable gives SHAW_ERR_ARG). For inversion loops that change
 * the material and for serving many source locations.
 *
 * The operators of `mat` on app's mesh into device buffers (current device):
 * coef_vp [nvp][4], coef_sp [nsp][2] (the layout of shaw_app_operators),
 * rho_inv [nvp]; any may be NULL. Equal to the host assembly bit for bit.  */
int shaw_app_assemble_device(shaw_app_t app, const shaw_material* mat, double* coef_vp, double* coef_sp,
                             double* rho_inv, void* stream);
/* shaw_forcing_table into a device buffer table [nsteps][m]; equal to the
 * host table up to the rounding of the device exp/sin (a few ulp).       */
int shaw_forcing_table_device(uint64_t m, const int* kinds, const double* delays, const double* periods,
                              double dt, uint64_t nsteps, double* table, void* stream);
/* gids[n] of n sources (depths_km, angles_deg) or, with depths_km NULL, n
 * receivers at the surface (angles_deg): shaw_map_source / shaw_map_receiver
 * for many locations in one launch, with the same distance formula and
 * tie rule (ties to the last gid).                                         */
int shaw_map_points_device(shaw_app_t app, uint64_t n, const double* depths_km, const double* angles_deg,
                           uint64_t* gids);

/* ------------------------------------------------------------------ FOM --
 * Full-order leapfrog on one GPU: x_v += dt J_vp x_s + dt rho^-1 f, then
 * x_s += dt J_sp x_v, for m forcing columns (m = 1 is the rank-1 state).   */
typedef struct shaw_fom* shaw_fom_t;

#define SHAW_FP64 0
#define SHAW_FP32 1
#define SHAW_KERNEL_AUTO 0   /* fused one-sweep step (ping-pong states) */
#define SHAW_KERNEL_SPLIT 1  /* velocity and stress half-step kernels, in place */
#define SHAW_KERNEL_FUSED 2

typedef struct {
  int precision;        /* SHAW_FP64 | SHAW_FP32 (state, coefficients, arithmetic) */
  int kernel;           /* SHAW_KERNEL_* */
  int rank1_forcing;    /* 1: add (dt*rho_inv)*f as KokkosBlas::mult (rank-1 form);
                           0: add (rho_inv*f)*dt as AddForcingRank2 (rank-2 form) */
  int cols_per_thread;  /* 0 = default; else 1, 2, 4 or 8 (must divide the column block) */
  int strip;            /* 0 = default; fused kernel: vp points per block row (incl. halo) */
  int segment;          /* 0 = default; fused kernel: radial rows per block */
  int prefetch_rows;    /* 0 = default; fused kernel: rows loaded ahead, 1 or 2 (the
                           multi-step kernels: 1; the one-wavefront-strip kernel at 3
                           or 4 steps: 1 or 2)                                        */
  int steps_per_launch; /* 0 = default: 6 on one-wavefront strips (cols_per_thread =
                           col_block <= 2, strip 64) when a launch fills the GPU, else 4
                           (loaded two rows ahead on those strips); fused kernel: 1..7
                           leapfrog steps per launch, the intermediate levels kept on
                           chip (strip >= 2 x steps; 5 to 7 only on one-wavefront
                           strips); a pack with a snapshot due at an intermediate step
                           falls back to fewer steps per launch; the last steps of a
                           run (up to two 7-step launches on one-wavefront strips)
                           go in as few, evenly split launches                         */
  int graph_steps;      /* 0 = default (64, rounded up to an even number of launches);
                           < 0: no graphs. Fused kernel without snapshots: whole chunks
                           of this many steps replay one captured HIP graph (kernels
                           read their step from a device counter); a multiple of
                           2 x steps_per_launch                                       */
  int col_block;        /* 0 = default (2); columns stored together per grid point
                           (power of two, at most 256 bytes of state): smaller blocks
                           give more, smaller workgroups                               */
} shaw_fom_options;

/* app: operators; m: forcing columns; dt: time step. opts may be NULL. */
int shaw_fom_create(shaw_app_t app, uint64_t m, double dt, const shaw_fom_options* opts,
                    shaw_fom_t* out);
/* vp gid of each column's point source (RankTwoForcing / RankOneForcing). */
int shaw_fom_set_sources(shaw_fom_t h, const uint64_t* vp_gid);
/* Forcing values [nsteps][m] for steps 1..nsteps (see shaw_forcing_table);
 * uploaded once, read by the velocity kernels.                             */
int shaw_fom_set_forcing(shaw_fom_t h, uint64_t nsteps, const double* table);
/* The forcing table of m signals (kinds/delays/periods as shaw_forcing_table)
 * generated on the device instead of uploaded (no counterpart in the
 * reference, which evaluates each forcing on the host every step).       */
int shaw_fom_set_signals(shaw_fom_t h, const int* kinds, const double* delays, const double* periods,
                         uint64_t nsteps);
/* Replace the FOM's operators by those of another material on the same mesh,
 * assembled on the device (shaw_app_assemble_device); the state is kept.
 * Equal to a FOM created on an app with that material.                    */
int shaw_fom_set_material(shaw_fom_t h, shaw_app_t app, const shaw_material* mat);
/* Record x_v at r receivers every `every` steps (step % every == 0), up to
 * `max_samples` samples; a device buffer LayoutLeft [r x T x m].          */
int shaw_fom_set_receivers(shaw_fom_t h, uint64_t r, const uint64_t* vp_gid, uint64_t every,
                           uint64_t max_samples);
/* Snapshot x_v (dof 0) or x_s (dof 1) every `every` steps for the listed
 * forcing columns into a device buffer LayoutLeft [n x T x ncols].        */
int shaw_fom_set_snapshots(shaw_fom_t h, int dof, uint64_t every, uint64_t ncols,
                           const uint64_t* cols, uint64_t max_samples);
/* Zero the state and the step/sample counters (fom_run.hpp:76-77). */
int shaw_fom_reset(shaw_fom_t h);
/* Enqueue nsteps leapfrog steps on the handle's stream. */
int shaw_fom_advance(shaw_fom_t h, uint64_t nsteps);
/* Wait for the handle's stream. */
int shaw_fom_sync(shaw_fom_t h);
/* Steps taken since reset. */
int shaw_fom_step(shaw_fom_t h, uint64_t* step);
/* Copy out the seismogram LayoutLeft [r x T x m] (T = samples so far). */
int shaw_fom_get_seismogram(shaw_fom_t h, double* out, uint64_t* samples);
int shaw_fom_get_snapshots(shaw_fom_t h, int dof, double* out, uint64_t* samples);
/* Current state, C order [nvp][m] and [nsp][m] (either may be NULL). */
int shaw_fom_get_state(shaw_fom_t h, double* vp, double* sp);
int shaw_fom_set_state(shaw_fom_t h, const double* vp, const double* sp);
/* Time `nsteps` steps with HIP events on the handle's stream: total ms and
 * the average duration of one leapfrog step (all its kernels).            */
int shaw_fom_time(shaw_fom_t h, uint64_t nsteps, double* total_ms, double* ms_per_step);
/* Enqueue nsteps steps in chunks of `chunk` steps with a HIP event between
 * chunks (no host wait between them), wait, and write each chunk's duration
 * [ms] to chunk_ms (ceil(nsteps / chunk) values). Steps run fused in
 * multi-step launches, so a chunk's duration over its steps is the per-step
 * time whose minimum and maximum runFom's perfTimes keep per step
 * (src/kokkos/fom_run.hpp:144-148).                                        */
int shaw_fom_time_chunks(shaw_fom_t h, uint64_t nsteps, uint64_t chunk, double* chunk_ms);
/* The reference's byte/flop model of one FOM step, as runFom prints it
 * (complexityFom, src/kokkos/fom_complexities.hpp:57-137, over
 * src/shared/complexities/{spmv,spmm,mult}.hpp): nnz of J_vp and J_sp, FP64
 * values and 8-byte CSR ordinals; rank2 = 0: spmv + mult_beta_one + spmv
 * (one forcing), rank2 = 1: spmm + the per-forcing add + spmm over m
 * columns. Host arithmetic only.                                           */
int shaw_fom_complexity(uint64_t nvp, uint64_t nsp, uint64_t nnz_vp, uint64_t nnz_sp, uint64_t m, int rank2,
                        double* mem_mb, double* flops);
/* Algorithmic HBM bytes of one leapfrog step of the selected kernel (the
 * state read and written once per launch, the coefficients read once per
 * launch, divided by the steps per launch), and the reference model
 * s*m*3*(nvp+nsp) + 8*(4*nvp + 2*nsp).                                    */
int shaw_fom_bytes_per_step(shaw_fom_t h, double* kernel_bytes, double* model_bytes);
/* Kernel configuration actually used: kernel, cols_per_thread, strip, segment,
 * prefetch_rows, steps_per_launch, col_block (any pointer may be null). */
int shaw_fom_config(shaw_fom_t h, int* kernel, int* cols_per_thread, int* strip, int* segment,
                    int* prefetch_rows, int* steps_per_launch, int* col_block);
/* Tile autotuner of the fused kernel (no counterpart in the reference, whose
 * launch shapes are fixed Kokkos policies): times every valid (columns per
 * thread, strip, segment, rows prefetched, steps per launch) for
 * trial_steps launches on the
 * current state, keeps the fastest (read it back with shaw_fom_config) and
 * restores the state. Needs the forcing table to cover the next step.     */
int shaw_fom_autotune(shaw_fom_t h, uint64_t trial_steps, double* best_ms_per_step);
/* POD basis of the device snapshot buffer of `dof` (all configured columns,
 * all samples: LayoutLeft [n x (T * ncols)]) without a host round trip, as
 * shaw_pod_basis; the snapshot buffer is consumed. FP64 FOM only.         */
int shaw_fom_pod(shaw_fom_t h, int dof, uint64_t k, double* phi, double* sigma);
/* POD in the inner product diag(w) (w [n] > 0): phi are w-orthonormal modes,
 * the left singular vectors of diag(sqrt w) S unscaled. No counterpart in the
 * reference (its POD, computeThinSVD, is Euclidean); used with
 * shaw_rom_project_energy and shaw_app energy weights.                     */
int shaw_fom_pod_weighted(shaw_fom_t h, int dof, uint64_t k, const double* w, double* phi,
                          double* sigma);
void shaw_fom_destroy(shaw_fom_t h);

/* ------------------------------------------------------------------ ROM --
 * Galerkin ROM. Projection in FP64 on the GPU; the reduced leapfrog for m
 * forcing columns in FP64 or BF16x3 split precision with FP32 accumulation
 * on MFMA, FP64 state accumulation.                                       */
typedef struct shaw_rom* shaw_rom_t;

#define SHAW_ROM_FP64 0
#define SHAW_ROM_BF16X3 1
/* BF16x4: the operator in three bf16 parts (hi, lo, residual), products
 * hi*q_hi + hi*q_lo + lo*q_hi + residual*q_hi with FP32 accumulation; no
 * counterpart in the reference (its ROM is FP64).                          */
#define SHAW_ROM_BF16X4 2

/* Ar_v = Phi_v^T (J_vp Phi_s) [kv x ks], Ar_s = Phi_s^T (J_sp Phi_v) [ks x kv],
 * row-major outputs; phi_v LayoutLeft [nvp x kv], phi_s LayoutLeft [nsp x ks]
 * (the lsv_* file layout).                                                */
int shaw_rom_project(shaw_app_t app, const double* phi_v, uint64_t kv, const double* phi_s,
                     uint64_t ks, double* ar_v, double* ar_s);
/* Energy-consistent projection (no counterpart in the reference, whose
 * Galerkin operators rom_compute_jacobians.hpp:58-109 are Euclidean):
 * Ar_v = Phi_v^T diag(mv) J_vp Phi_s and Ar_s = -Ar_v^T, for bases that are
 * orthonormal in the FOM energy weights (mv on velocity points; stress bases
 * orthonormal in the stress weights). The FOM satisfies
 * Mv J_vp = -(Ms J_sp)^T on every live point, so Ar_s is the Galerkin
 * operator itself; stated exactly, the reduced leapfrog conserves a reduced
 * energy and operator rounding (BF16x3) cannot excite growing modes.       */
int shaw_rom_project_energy(shaw_app_t app, const double* phi_v, uint64_t kv, const double* phi_s,
                            uint64_t ks, const double* mv, double* ar_v, double* ar_s);

/* shaw_rom_project on device pointers that allocates nothing. No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops that forbid
 * hipMalloc). coef_vp [nvp][4] and coef_sp [nsp][2] (as shaw_app_operators
 * returns them), phi_v LayoutLeft [nvp x kv], phi_s LayoutLeft [nsp x ks],
 * ar_v [kv][ks] and ar_s [ks][kv]: all device memory of the current device;
 * stream: a hipStream_t (NULL: the null stream). Only enqueues: the
 * operators are ready when the stream reaches them. The split-K partial
 * tiles live in a 16 MiB array of the library on each GPU, so kv * ks is at
 * most 2097152 (SHAW_ERR_ARG otherwise, before any launch), and calls that
 * overlap on one GPU race. Same kernels and chunking as shaw_rom_project:
 * equal operators bit for bit.                                             */
int shaw_rom_project_noalloc(uint64_t nr, uint64_t nth, const double* coef_vp, const double* coef_sp,
                             const double* phi_v, uint64_t kv, const double* phi_s, uint64_t ks,
                             double* ar_v, double* ar_s, void* stream);

/* The basis Phi LayoutLeft [n x k] made orthonormal in the inner product
 * diag(w) (w [n] > 0) without changing its span: phi_out = Phi R^-1 with
 * R^T R = Phi^T diag(w) Phi (R upper triangular, Cholesky), so
 * phi_out^T diag(w) phi_out = I and coordinates in phi_out convert back to
 * Phi's by q = R^-1 q_out. r_out [k][k] row-major (may be NULL). With the
 * FOM energy weights this turns the reference's POD bases (lsv files) into
 * bases for shaw_rom_project_energy. No counterpart in the reference.     */
int shaw_basis_orthonormalize(const double* phi, uint64_t n, uint64_t k, const double* w, double* phi_out,
                              double* r_out);

/* Device-resident POD bases. The reference's offline chain moves each basis
 * through files (computeThinSVD writes lsv_<dof>, rom_load_basis.hpp:56-101
 * reads it back, rom_compute_jacobians.hpp:58-109 projects it); here the
 * bases of both dofs stay in HBM from the POD to the projection, and only
 * what the ROM needs (Ar, the rows at source and receiver gids) comes to
 * the host. At config 3's size a basis pair is 26 GB, which a host round
 * trip moves twice over PCIe.
 * shaw_fom_pod_basis: kv and ks modes of the FOM handle's snapshot buffers
 * (both consumed), Euclidean as shaw_fom_pod, or with wv [nvp], ws [nsp]
 * (both or neither) as shaw_fom_pod_weighted; equal to those bit for bit.
 * The first j columns of a basis are its j-mode basis (POD bases nest). */
typedef struct shaw_basis* shaw_basis_t;
int shaw_fom_pod_basis(shaw_fom_t h, uint64_t kv, uint64_t ks, const double* wv, const double* ws,
                       shaw_basis_t* out);
int shaw_basis_info(shaw_basis_t b, uint64_t* n_vp, uint64_t* n_sp, uint64_t* kv, uint64_t* ks,
                    uint64_t* nsnap_vp, uint64_t* nsnap_sp);
/* the snapshot matrix's singular values of dof (0 vp, 1 sp): nsnap values */
int shaw_basis_sigma(shaw_basis_t b, int dof, double* sigma);
/* out [m][k] row-major: the first k modes at grid points gids[m] (the rows
 * rom_problem_rank_one.hpp:133-147 and main_reconstruct_seismo.cc:13-41
 * take from Phi_v for the forcing vector and the receivers)               */
int shaw_basis_rows(shaw_basis_t b, int dof, const uint64_t* gids, uint64_t m, uint64_t k, double* out);
/* phi LayoutLeft [n x k]: the first k modes (the lsv_<dof> layout)       */
int shaw_basis_download(shaw_basis_t b, int dof, uint64_t k, double* phi);
/* the operators of the first kv / ks modes: mv NULL, as shaw_rom_project;
 * else the energy pair of shaw_rom_project_energy. Equal bit for bit to
 * those calls on the downloaded bases.                                    */
int shaw_basis_project(shaw_app_t app, shaw_basis_t b, uint64_t kv, uint64_t ks, const double* mv,
                       double* ar_v, double* ar_s);
void shaw_basis_destroy(shaw_basis_t b);

typedef struct {
  uint64_t kv, ks;
  double dt;
  int precision;         /* SHAW_ROM_FP64 | SHAW_ROM_BF16X3 | SHAW_ROM_BF16X4 */
  const double* ar_v;    /* [kv][ks] */
  const double* ar_s;    /* [ks][kv] */
  uint64_t r;            /* receivers */
  const double* c;       /* [r][kv] = rows of Phi_v at the receiver gids */
  uint64_t max_cols;     /* resident column slots */
  uint64_t max_samples;  /* seismogram samples per column */
} shaw_rom_desc;

int shaw_rom_create(const shaw_rom_desc* d, shaw_rom_t* out);
/* Admit n columns into free slots: b [kv][n] (rho_inv[g] Phi_v[g,:] per
 * column), forcing values [nsteps][n], horizons (steps per column), sample
 * period `every`. slots_out[n] receives the slot ids. Fails with
 * SHAW_ERR_MEMORY when not enough slots are free. Waits for the engine's
 * launches in flight (they write back every slot of their workgroups).    */
int shaw_rom_admit(shaw_rom_t h, uint64_t n, const double* b, uint64_t nsteps, const double* ftab,
                   const uint64_t* horizons, uint64_t every, uint64_t* slots_out);
/* shaw_rom_admit with each column's forcing series evaluated on the device
 * from its signal (kinds/delays/periods [n], shaw_forcing_table's
 * conventions, time step the engine's dt) instead of a host table.        */
int shaw_rom_admit_signals(shaw_rom_t h, uint64_t n, const double* b, const int* kinds, const double* delays,
                           const double* periods, uint64_t nsteps, const uint64_t* horizons, uint64_t every,
                           uint64_t* slots_out);
/* Macro steps (BF16 engines; no counterpart in the reference): with m in
 * 2..64 a tile of 16 column slots whose active columns are at steps
 * congruent mod m and sampled at multiples of m (or not at all) advances,
 * from the first multiple of m at which all of them are quiet (their forcing
 * at or below 2^-64 of its peak from there on, found at admission), m steps
 * per product by G = (F^m - I) / dt, F the leapfrog's one-step matrix of
 * [q_v; q_s] (built on the host from the engine's FP64 operators by
 * stepping the identity m times), while every column stays at least m steps
 * from its horizon; all other steps are one-steps, in the same launch. The
 * results differ from one-step BF16 runs at the BF16 rounding level, and a
 * column's steps depend on the other columns of its tile. m = 0 (the
 * default) turns them off.                                                */
int shaw_rom_set_macro_steps(shaw_rom_t h, uint32_t m);
/* Time budget for engines with macro steps (no counterpart in the
 * reference; for serving, where a launch lasts as long as its slowest tile):
 * with w in 1..64, shaw_rom_advance(h, n) gives every tile the time of n
 * one-steps, a macro step counting as w of them, instead of n steps; a tile
 * that takes macro steps then advances up to about n m / w steps (columns
 * stop at their horizon). How far each column got is known once the
 * engine's work is done: shaw_rom_steps, shaw_rom_state and the sample
 * counts of retirement read it back. w = 0 (the default) turns it off.      */
int shaw_rom_set_time_budget(shaw_rom_t h, uint32_t w);
/* Step count of each of n admitted slots, after the engine's enqueued work
 * (waits for it).                                                          */
int shaw_rom_steps(shaw_rom_t h, uint64_t n, const uint64_t* slots, uint64_t* steps);
/* Advance every active column by up to nsteps (columns stop at their horizon;
 * under a time budget, see shaw_rom_set_time_budget).                      */
int shaw_rom_advance(shaw_rom_t h, uint64_t nsteps);
int shaw_rom_sync(shaw_rom_t h);
/* Copy a column's seismogram LayoutLeft [r x T] and final reduced states
 * (any output may be NULL), then free the slot.                            */
int shaw_rom_retire(shaw_rom_t h, uint64_t slot, double* seis, double* qv, double* qs,
                    uint64_t* samples);
/* Retire n slots at once (one stream sync and one staged copy): seismograms
 * into seis [n][r x max_samples] (each LayoutLeft [r x T_j], T_j samples
 * stored, the rest untouched), T_j into samples (may be NULL). No counterpart
 * in the reference; for serving, where a request retires many columns.     */
int shaw_rom_retire_batch(shaw_rom_t h, uint64_t n, const uint64_t* slots, uint64_t max_samples,
                          double* seis, uint64_t* samples);
/* Current reduced states of an admitted column (kv and ks values) and its
 * step count; any output may be NULL.                                      */
int shaw_rom_state(shaw_rom_t h, uint64_t slot, double* qv, double* qs, uint64_t* step);
int shaw_rom_free_slots(shaw_rom_t h, uint64_t* n);
/* Non-blocking: *idle = 1 when the engine's stream has no work in flight. */
int shaw_rom_query(shaw_rom_t h, int* idle);
int shaw_rom_time(shaw_rom_t h, uint64_t nsteps, double* total_ms);
void shaw_rom_destroy(shaw_rom_t h);

/* Shape of an engine: kv, ks, receivers, column slots, samples per column
 * (any pointer may be NULL).                                               */
int shaw_rom_info(shaw_rom_t h, uint64_t* kv, uint64_t* ks, uint64_t* r, uint64_t* max_cols,
                  uint64_t* max_samples);

/* ------------------------------------------------------------ serving --
 * Continuous batching of ROM requests on resident engines. No counterpart in
 * the reference, which runs one ROM problem per process
 * (src/kokkos/main.cc:66-73, rom_problem_rank_one.hpp:123-159).
 * An engine is a table of the shaw_rom_* calls the loop makes, so the same
 * loop drives GPU engines (shaw_engine_from_rom) and any other engine with
 * the same contract (a CPU one in the tests).                              */
typedef struct {
  void* ctx;
  uint64_t kv, ks, r, capacity;  /* modes, receivers, column slots */
  int (*admit)(void* ctx, uint64_t n, const double* b, uint64_t nsteps, const double* ftab,
               const uint64_t* horizons, uint64_t every, uint64_t* slots);
  int (*advance)(void* ctx, uint64_t nsteps);  /* enqueue only */
  int (*sync)(void* ctx);
  int (*retire_batch)(void* ctx, uint64_t n, const uint64_t* slots, uint64_t max_samples, double* seis,
                      uint64_t* samples);
  int (*free_slots)(void* ctx, uint64_t* n);
  int (*query)(void* ctx, int* idle);  /* non-blocking: *idle = 1 when all enqueued work is done */
  /* admission with the forcing series evaluated by the engine from signal
   * parameters (shaw_rom_admit_signals); NULL: requests need a table      */
  int (*admit_signals)(void* ctx, uint64_t n, const double* b, const int* kinds, const double* delays,
                       const double* periods, uint64_t nsteps, const uint64_t* horizons, uint64_t every,
                       uint64_t* slots);
  /* step count of each slot once the enqueued work is done (shaw_rom_steps);
   * NULL: every active column advances by exactly the steps asked for      */
  int (*steps)(void* ctx, uint64_t n, const uint64_t* slots, uint64_t* steps);
} shaw_engine;
/* The table of a GPU engine: shaw_rom_admit/advance/sync/retire_batch/free_slots/
 * query/admit_signals/steps. */
int shaw_engine_from_rom(shaw_rom_t rom, shaw_engine* out);

typedef struct {
  uint64_t id;
  int engine;            /* index into the engines */
  uint64_t m;            /* forcing columns */
  uint64_t horizon;      /* steps */
  uint64_t every;        /* seismogram sampling period (0: none) */
  double arrival_s;      /* arrival time on the server clock (0 = start of the run) */
  const double* b;       /* [kv][m]: reduced forcing vector of each column */
  const double* ftab;    /* [horizon][m]: forcing values */
  double* seis;          /* out: [m][r x horizon/every] (NULL: discarded) */
  double admitted_s;     /* out: admission time */
  double done_s;         /* out: completion time (latency = done_s - arrival_s) */
  /* instead of ftab (NULL): each column's signal, kinds/delays/periods [m]
   * as shaw_forcing_table, evaluated by the engine (admit_signals)        */
  const int* kinds;
  const double* delays;
  const double* periods;
} shaw_request;

#define SHAW_SERVE_FIFO 0        /* arrival order; the head is never passed */
#define SHAW_SERVE_LEAST_WORK 1  /* least horizon x columns first among the requests that fit */

typedef struct {
  uint64_t hbm_budget;   /* bytes of resident columns (states, forcing, seismograms) */
  uint64_t chunk_steps;  /* steps per engine launch between admissions */
  int policy;            /* SHAW_SERVE_* */
  double max_wait_s;     /* LEAST_WORK: a request waiting this long is admitted before any
                            other and no request for its engine passes it (none at all when
                            it waits for HBM budget; 0: no bound) */
  double wall_limit_s;   /* fail with SHAW_ERR_STATE past this (0: none) */
  /* optional clock [s] and idle wait (tests): NULL = steady clock and sleep */
  double (*clock)(void* clock_ctx);
  void (*wait_until)(void* clock_ctx, double t);
  void* clock_ctx;
} shaw_serve_options;

typedef struct {
  uint64_t requests, forcing_steps, launches, max_queue;
  double span_s;         /* first arrival to last completion */
  double p50_latency_s, p99_latency_s, max_latency_s;
  double max_wait_s;     /* longest time from arrival to admission */
  double throughput;     /* forcing-steps / span_s */
} shaw_serve_stats;

/* Serve n requests arriving at their arrival times until all are done;
 * writes each request's seismograms, admitted_s and done_s, and the stats.
 * Engines run independently: an engine that finished its chunk retires its
 * finished requests, admits queued ones and advances again while the others
 * are still running.                                                       */
int shaw_serve(const shaw_engine* engines, int n_engines, const shaw_serve_options* opts,
               shaw_request* reqs, uint64_t n, shaw_serve_stats* stats);

/* POD basis: the k leading left singular vectors of the snapshot matrix
 * S LayoutLeft [n x ns] (the lsv_<dof> file content) and ns singular values,
 * descending, in FP64 on the GPU by a deflated method of snapshots: each pass
 * eigen-decomposes the Gram matrix of the residual snapshots (Householder
 * tridiagonalization + implicit symmetric QR, on the host), keeps
 * the modes it resolves (sigma > 1e-3 of the residual's largest), orthogonal-
 * izes them (Cholesky) and projects them out of the residual. Fails if k
 * exceeds the numerical rank. Replaces computeThinSVD
 * (src/tools/main_eigen_svd.cc:113-197). phi: LayoutLeft [n x k].          */
int shaw_pod_basis(const double* s, uint64_t n, uint64_t ns, uint64_t k, double* phi, double* sigma);
/* As shaw_pod_basis, but computes up to kmax modes and stops at the numerical
 * rank (singular values above 1e-14 sigma_1); *k receives the count. This is
 * what computeThinSVD's thin U (src/tools/main_eigen_svd.cc:129-133) holds
 * for a full-rank snapshot matrix; for a rank-deficient one only the resolved
 * columns are written (the trailing columns of a thin SVD there are an
 * arbitrary orthonormal completion). phi: [n x kmax] capacity.             */
int shaw_pod_basis_upto(const double* s, uint64_t n, uint64_t ns, uint64_t kmax, double* phi, double* sigma,
                        uint64_t* k);
/* out = phi q: the FOM state (or, with the rows of phi at receiver gids, the
 * seismogram) of reduced coordinates, on the GPU in FP64. phi LayoutLeft
 * [n x k], q LayoutLeft [k x ncols], out LayoutLeft [n x ncols]; each entry
 * is the sum over modes in index order without FMA, as the loops of
 * src/tools/main_reconstruct_seismo.cc:28-38 (reconstructSeismogram) and the
 * gemv of src/tools/main_reconstruct_fom_state.cc:158-160 (reconstructFomState). */
int shaw_reconstruct(const double* phi, uint64_t n, uint64_t k, const double* q, uint64_t ncols, double* out);

/* ---------------------------------------------------------- diagnostics --
 * One MFMA instruction on one wavefront with lane-wise operands, for
 * checking the operand/result lane layouts the ROM kernels assume.
 * kind 0: v_mfma_f64_16x16x4f64, a[64], b[64] -> d[64][4];
 * kind 1: v_mfma_f32_16x16x16bf16_1k, a[64][4], b[64][4] (rounded to bf16)
 *         -> d[64][4];
 * kind 2: v_mfma_f32_16x16x32_bf16, a[64][8], b[64][8] (rounded to bf16)
 *         -> d[64][4]. No reference counterpart.                          */
int shaw_debug_mfma(int kind, const double* a, const double* b, double* d);
/* Host symmetric eigen-decomposition used by the POD (Householder
 * tridiagonalization + implicit symmetric QR with Wilkinson shifts, Golub &
 * Van Loan, Matrix Computations, 4th ed., Algorithms 8.3.1-8.3.3): a row-major [n x n] symmetric,
 * w[n] eigenvalues (unordered), v row-major [n x n] with eigenvectors in its
 * columns. CPU only. No reference counterpart.                             */
int shaw_debug_symeig(int n, const double* a, double* w, double* v);
/* Time `reps` launches of a plain streaming copy kernel (k_copy) of `bytes`
 * bytes (read `bytes`, write `bytes`); used to calibrate the HBM traffic
 * counters and to state the achievable copy bandwidth. No reference
 * counterpart.                                                             */
int shaw_debug_copy(uint64_t bytes, int reps, double* ms_per_copy);

#ifdef __cplusplus
}
#endif
#endif /* SHAW_AMD_H_ */
