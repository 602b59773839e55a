/*
 * btpm.h -- C ABI of libbtpm, the MI355X (gfx950) Thermal stage of the
 * binary-asteroid ATPM: per-column 1-D explicit conduction with a Newton
 * surface energy balance, facet self-heating, crater sub-facet scattering
 * and re-emission, rotation-to-rotation convergence with per-column freezing.
 *
 * Plain pointers and sizes only.  Host arrays are read during the call that
 * receives them; the library owns every device buffer.  All functions return
 * BTPM_OK (0) or a negative BTPM_ERR_* code; btpm_last_error() gives the
 * message of the last failure on the calling thread.
 *
 * Reference interfaces replaced (paths relative to the reference tree):
 *   - the in-process smooth path of `single_model`:
 *     initialisetemperatures / inertiadifference / instequilibrium /
 *     shape_converged, 175706/Run-04/thermal.cpp:1861-2033, 2422-2454,
 *     3590-3607 (no ABI in the reference; these entry points add one);
 *   - the crater path behind the extern "C" CUDA host API,
 *     175706/Run-04/thermal_GPU.cu:698-1279 (also available symbol for
 *     symbol through btpm_compat.h);
 *   - the host loops load_GPU_roughness_globalillumination /
 *     _globalheating, 175706/Run-04/thermal.cpp:2693-2803;
 *   - singlescattering / multiplescattering, thermal.cpp:1745-1855.
 */
#ifndef BTPM_H
#define BTPM_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BTPM_OK 0
#define BTPM_ERR_ARG (-1)     /* invalid argument or size                 */
#define BTPM_ERR_HIP (-2)     /* HIP runtime failure                      */
#define BTPM_ERR_NOMEM (-3)   /* device memory exhausted                  */
#define BTPM_ERR_STATE (-4)   /* call not valid in the object's state     */

#define BTPM_NS 100           /* crater sub-facets per parent facet       */
#define BTPM_NX 56            /* depth nodes below the surface            */

typedef struct btpm_ctx btpm_ctx;
typedef struct btpm_shape btpm_shape;
typedef struct btpm_solve btpm_solve;
typedef struct btpm_server btpm_server;

/* Model constants.  Field meanings follow the reference globals of
 * 175706/Run-04/thermal.cpp:14-83 (NTAU, NX, rotation_period, EPS, SIGMA,
 * IR_ABS, TACC); newton_tol is TACC/10 in the reference (:1936), max_rev the
 * 301-rotation cap (:3592) and check_after the `rev > 5` gate (:3600).
 * newton_cap bounds the Newton loop, which the reference leaves unbounded.  */
typedef struct {
    int ntau;
    int nx;              /* must be BTPM_NX */
    float rot_period;    /* s */
    float eps;
    float sigma;
    float ir_abs;
    float tacc;          /* K */
    float newton_tol;    /* K */
    int max_rev;
    int check_after;
    int newton_cap;
} btpm_params;

/* A body: its facets' illumination and the sparse facet view factors.
 * illum is [nf][ntau] (facet-major, the reference's totalillumination).
 * If scatter != 0, illum holds the DIRECT illumination solar*cos(i)*shadow
 * and the library applies singlescattering (+ multiplescattering when
 * scatter == 2) on the GPU with scatter_acc (W m^-2).
 * CSR rows are receiving facets: row f lists (colidx[k], vf[k]) = (g, VF_fg)
 * as in <shape>_selfheating_list.txt (T3905/main.cpp:418-434).              */
typedef struct {
    int nf;
    const float *illum;
    const float *albedo;     /* [nf] */
    const float *numrevs;    /* [nf] revolutions per cycle, NULL = 1 */
    int nnz;
    const int *rowptr;       /* [nf+1] or NULL */
    const int *colidx;       /* [nnz] */
    const float *vf;         /* [nnz] */
    int scatter;             /* 0: illum is total; 1: single; 2: multiple */
    float scatter_acc;
} btpm_body;

/* The crater (roughness) model attached to every facet of a body.
 * vf[i*100+j] = VF(j->i), the transposed storage that
 * initialise_GPU_constants receives (T3905/main.cpp:1994-2000).
 * Crater illumination is either given as `illum` [nf][ntau][100] (total,
 * after scattering) or computed on the GPU from shadow/cosang
 * [nf][ntau][100], `solar` and the shape->crater global illumination
 * (roughness_single/multiple_scattering, thermal_GPU.cu:241-341).
 * drf [nnz][100] (optional) are the shape->crater view factors per CSR entry
 * of the body (DepRoughFacet::viewfactor(1..100), thermal.cpp:318-334); they
 * drive RGI and RGH.
 * A third source of direct crater sunlight (no [nf][ntau][100] arrays at
 * all): sun_local [nf][ntau][3] is the Sun direction in each facet's crater
 * frame (z = facet normal) with length = the facet's own sunlit fraction
 * (0 when shadowed or eclipsed); vis_table [n_el][n_az][100] the sunlit
 * fraction of every sub-facet at elevation i*90/(n_el-1) deg and azimuth
 * i*360/n_az deg, interpolated bilinearly; normals [100][3] the sub-facet
 * normals.  The GPU evaluates shadow = |d| vis(d/|d|) and
 * cosang = max(0, n_j . d/|d|) inside the scattering kernel (the role of
 * sfcpartialshadow, 3905/Thermal/NominalPole/GPU_code.cu:152-210, with a
 * tabulated cap instead of its hemisphere rule).                            */
typedef struct {
    const float *vf;
    const float *illum;
    const float *shadow;
    const float *cosang;
    float solar;
    int multiple;            /* multiple scattering */
    float scatter_acc;
    const float *drf;
    const float *sun_local;
    const float *vis_table;
    int n_el, n_az;
    const float *normals;
} btpm_crater;

typedef struct {
    int rotations;               /* smooth rotations executed            */
    int converged;               /* smooth: every facet froze            */
    long long column_steps;      /* smooth: integrated column-steps      */
    int crater_rotations;
    int crater_converged;
    long long crater_column_steps;
    double latency_s;            /* admission to completion (batch/server) */
} btpm_solve_info;

typedef struct {
    int rotations_per_sync;      /* rotations enqueued between host checks;
                                    0: chosen by the scheduler from the
                                    measured time per rotation             */
    int use_graph;               /* capture each rotation set in a hipGraph */
    int max_active_craters;      /* crater solves resident at once (0: by memory) */
    double mem_fraction;         /* admission keeps (1 - mem_fraction) of
                                    the HBM free at start unused; it checks
                                    live free memory, so results of finished
                                    solves count until they are destroyed  */
    int profile;                 /* time every conduction launch with events */
    int autotune;                /* choose tile widths per size class by timed
                                    trial rotations (results are identical for
                                    every choice); decisions are kept in the
                                    context.  Default 1                        */
    int fast_extra;              /* extra rotations, after every rotation of
                                    all sets, for sets at most 1/8 the columns
                                    of the largest resident set (they finish
                                    in fewer shared rotations; results do not
                                    change).  In profiled runs they are not
                                    timed.  Default 3                          */
} btpm_batch_opts;

typedef struct {
    double wall_s;
    int solves;
    int converged;
    long long column_steps;      /* integrated (executed) column-steps     */
    long long column_steps_all;  /* including frozen columns              */
    int graph_launches;
    int syncs;
    /* with opts.profile: summed launch durations and launch counts of the
     * two conduction kernels, and the column-steps each integrated        */
    double smooth_kernel_s;
    double crater_kernel_s;
    int smooth_launches;
    int crater_launches;
    long long smooth_column_steps;
    long long crater_column_steps;
    int max_resident;            /* most solves resident at once           */
    int deferrals;               /* set builds that met an out-of-memory and
                                    were retried after other solves ran   */
    int tune_rotations;          /* rotations run as timed tuning trials    */
    int tune_decisions;          /* size classes the autotuner decided      */
    long long fast_rotations;    /* extra fast-lane rotations enqueued      */
    int mem_waits;               /* solves whose admission waited for device
                                    memory (the HBM-footprint budget)      */
    /* with opts.profile: the facet self-heating phase (one launch per shape
     * with view factors), its summed duration, launches and the smooth sets
     * those launches covered (set-rotations)                                */
    double heat_kernel_s;
    int heat_launches;
    long long heat_set_rotations;
} btpm_batch_stats;

/* Illumination geometry and shadows of a (binary) shape over one cycle: the
 * producer of the direct facet illumination the solves consume
 * (asteroidgeometry, astshadowtest / astshadow and the direct term of
 * singlescattering, T175706/thermal.cpp:1272-1471, 1612-1707, 1757; the
 * reference runs them on the host).  Vertices equpos [nv][3] in km in each
 * body's equatorial frame (the secondary's include its orbit), the first
 * nv_primary on the primary; facets [nf][3]; each body turns rev_primary /
 * rev_secondary times per cycle about its z axis; pol = (r in AU, ecliptic
 * latitude, longitude) of the system and rot = (pole ecliptic latitude,
 * longitude) in radians; au_km = 149.6e6 in the reference.                  */
typedef struct {
    int nv, nf, ntau, nv_primary;
    const double *equpos;
    const int *facets;
    double rev_primary, rev_secondary;
    double pol[3];
    double rot[2];
    double au_km;
    double sun_scale;    /* 0: sun_facet holds unit directions; > 0: sun_scale
                            times the direction, as the reference's
                            sfcillumvector (10000, thermal.cpp:1453-1455)   */
} btpm_geometry;

/* ---- library ------------------------------------------------------------ */
int btpm_version(void);
const char *btpm_last_error(void);
void btpm_params_default(btpm_params *p, int ntau, float rot_period, float tacc);
void btpm_batch_opts_default(btpm_batch_opts *o);

/* set_cuda_device (thermal_GPU.cu:698-702) + module state: one context per
 * device, each with its own stream.                                         */
int btpm_ctx_create(int device, btpm_ctx **out);
int btpm_ctx_destroy(btpm_ctx *ctx);
/* free_bytes counts the driver's free memory plus the device blocks the
 * library keeps cached for reuse (released buffers are not handed back to
 * the driver until the context is destroyed or an allocation fails).       */
int btpm_ctx_mem_info(btpm_ctx *ctx, size_t *free_bytes, size_t *total_bytes);
/* Caps the device memory the library holds on the context's device (live
 * and cached buffers of every context there) at `bytes`; 0 removes the cap.
 * An allocation beyond it fails as if the device were full, and batch /
 * server admission budgets against it.  The call returns the library's
 * cached (idle) blocks to the driver.  Lets a serving process share a GPU.
 * No counterpart in the reference (which allocates whatever its compile-time
 * sizes need).                                                              */
int btpm_ctx_set_mem_limit(btpm_ctx *ctx, size_t bytes);

/* ---- shapes (inputs shared by many solves) ---------------------------------
 * Uploads the body (and crater model), runs facet scattering, RGI and crater
 * scattering on the GPU.  Replaces initialise_GPU_constants,
 * initialise_GPU_*_arrays, load_GPU_globalillumination,
 * GPU_single_scattering / GPU_multiple_scattering
 * (thermal_GPU.cu:708-1172) and their host loops.                            */
int btpm_shape_create(btpm_ctx *ctx, const btpm_params *p, const btpm_body *body,
                      const btpm_crater *crater, btpm_shape **out);
int btpm_shape_destroy(btpm_shape *shape);
/* total facet illumination F after scattering, [nf][ntau] */
int btpm_shape_facet_illum(btpm_shape *shape, float *out);
/* total crater illumination after scattering, [nf][ntau][100]
 * (obtain_GPU_roughness_illumination, thermal_GPU.cu:1122-1135)            */
int btpm_shape_crater_illum(btpm_shape *shape, float *out);

/* ---- solves -------------------------------------------------------------
 * One thermal-inertia solve of a shape: the smooth facets to convergence
 * and, with with_craters, then RGH from the converged facets and the crater
 * columns to convergence (single_model's per-TI body,
 * thermal.cpp:3676-3760).  ti_per_facet (optional, [nf]) overrides ti.
 * Every solve starts from the reference initial state.                      */
int btpm_solve_create(btpm_ctx *ctx, btpm_shape *shape, float ti, const float *ti_per_facet,
                      int with_craters, btpm_solve **out);
int btpm_solve_destroy(btpm_solve *s);
/* Scales every absorbed flux of this solve: the facet illumination after
 * scattering and the crater illumination, both linear in the solar flux
 * (e.g. (r_ref / r)^2 places the shape at distance r instead of r_ref).
 * Call before the solve runs.  No counterpart in the reference, which bakes
 * the solar constant into its illumination (T3905/main.cpp:905); it lets
 * requests at different heliocentric distances share one resident shape.  */
int btpm_solve_set_flux_scale(btpm_solve *s, float scale);
/* Warm start: s starts from the final state of `from` (same shape) instead
 * of the reference initial state -- the carried state of the reference's TI
 * sweep, which initialises once and runs every later thermal inertia from
 * the previous one's converged temperatures (175706/Run-04/thermal.cpp:
 * 3554-3576, 3676-3700; 3905/Thermal/NominalPole/main.cpp:2478-2517).  The
 * smooth set takes from's final depth profiles and last-rotation surface
 * temperatures; the crater set takes from's crater state when both carry
 * craters (otherwise it starts as usual).  A `from` at TI 0 everywhere
 * carries no depth profile, so s then starts from the initial state, as the
 * reference initialises at its first TI > 0.  Convergence status restarts
 * (the rev > 5 gate counts from 0 again, reset_asteroid_status).
 * `from` must have finished, or run before s in the same batch or server
 * (s's smooth phase waits for from's smooth phase, its crater phase for the
 * whole of from).  Until s has taken the copy, from cannot be destroyed
 * (BTPM_ERR_STATE).  Call before s runs.                                   */
int btpm_solve_warm_start(btpm_solve *s, btpm_solve *from);
/* Runs one solve alone through the persistent kernels, checking
 * convergence on the host after every rotation (no batching, no graphs).   */
int btpm_solve_run(btpm_solve *s);
/* Runs one solve alone in the reference's GPU execution structure: one
 * finite-difference and one boundary-condition launch per time step with a
 * host synchronisation after each, the depth profile in global memory, the
 * crater self-heating as a per-lane loop, and a convergence kernel whose
 * flags the host reads every rotation (GPU_inertia_difference /
 * GPU_converged, T175706/thermal_GPU.cu:430-692, 1021-1026).  Same
 * algorithm and results to tolerance; the baseline the batched path is
 * measured against.  rotation_cap > 0 stops each phase after that many
 * rotations (timing samples); 0 runs to convergence or max_rev.            */
int btpm_solve_run_stepwise(btpm_solve *s, int rotation_cap);
int btpm_solve_get_info(btpm_solve *s, btpm_solve_info *info);
/* final surface temperatures: smooth [nf][ntau] (writeglobaltemperature's
 * values), crater [nf][ntau][100] (obtain_GPU_roughness_temperatures'
 * sfc_stemp1), depth profiles [nf][nx+1] / [nf][nx+1][100].  Surface
 * temperatures are floored at +0 K (a facet the Sun never reaches cools to
 * 0 K, where rounding can leave -1e-15 K or -0; a Planck function turns
 * those into large negative intensities).                                   */
int btpm_solve_surface_temps(btpm_solve *s, float *out);
int btpm_solve_crater_temps(btpm_solve *s, float *out);
int btpm_solve_profile(btpm_solve *s, float *out);
int btpm_solve_crater_profile(btpm_solve *s, float *out);

/* ---- batches ---------------------------------------------------------------
 * Runs n solves to convergence together: solves are admitted while HBM
 * allows, every rotation advances all resident sets in one launch per kernel
 * kind (captured in a hipGraph), convergence and phase changes are checked
 * on the host every rotations_per_sync rotations, and finished solves make
 * room for waiting ones.                                                    */
int btpm_batch_run(btpm_ctx *ctx, btpm_solve **solves, int n, const btpm_batch_opts *opts,
                   btpm_batch_stats *stats);

/* ---- server (continuous batching) -------------------------------------------
 * A request loop on its own host thread: submitted solves are admitted as
 * memory and slots free up, advanced together with the resident ones, and
 * completed with their latency.  Starting a server returns the device
 * blocks the library keeps cached (from earlier batches) to the driver.     */
int btpm_server_start(btpm_ctx *ctx, const btpm_batch_opts *opts, btpm_server **out);
int btpm_server_submit(btpm_server *srv, btpm_solve *s);
/* blocks until s completes (or timeout_s elapses: BTPM_ERR_STATE)          */
int btpm_server_wait(btpm_server *srv, btpm_solve *s, double timeout_s);
/* Blocks until one of solves[0..n) has completed (or timeout_s elapses:
 * BTPM_ERR_STATE) and stores its index in *index (the lowest index among
 * completed ones).  Lets a client consume results in completion order, so
 * finished solves do not hold device memory while it waits for earlier
 * requests.  No counterpart in the reference (which serves no requests).  */
int btpm_server_wait_any(btpm_server *srv, btpm_solve *const *solves, int n, double timeout_s,
                         int *index);
int btpm_server_stop(btpm_server *srv, btpm_batch_stats *stats);

/* ---- kernel-level entry points (tests and tools) ----------------------------
 * Crater self-heating product of the crater conduction kernel, standalone:
 * q[n][j] = sum_i VF(j->i) e[n][i] for n < ncol, computed with the same
 * fp16 hi/lo MFMA split; e is [ncol][100], q is [ncol][100].                */
int btpm_crater_contract(btpm_ctx *ctx, const float *vf, const float *e, int ncol, float *q);
/* The same product without any allocation, on device memory: vf [100*100]
 * (vf[i*100+j] = VF(j->i)), e [ncol][100] and q [ncol][100] are device
 * pointers; the packed fp16 hi/lo tiles of VF go to a fixed 50,432-byte
 * array in device memory (one per device, held until the process ends)
 * instead of a buffer the library allocates.  stream is a hipStream_t: the
 * call is enqueued on it and returns at once (graph capture allowed); NULL
 * runs on the context's stream and waits.  Results equal
 * btpm_crater_contract's bit for bit.  Calls that overlap on one device
 * race on the array (calls on one stream are ordered).  ncol at most
 * BTPM_NOALLOC_MAX_COLS.  No counterpart in the reference; for callers that
 * cannot allocate device memory (code under HIP graph capture, serving
 * loops that must not allocate).                                           */
#define BTPM_NOALLOC_MAX_COLS (1 << 24)
int btpm_crater_contract_noalloc(btpm_ctx *ctx, const float *vf, const float *e, int ncol,
                                 float *q, void *stream);

/* Outputs [nf][ntau] (host memory): shadow (1 lit, 0 back-facing or shadowed
 * by a facet nearer the Sun, including the partner body: eclipses), illum =
 * solar cos(angle) shadow in W/m^2 (may be NULL), sun_facet [nf][ntau][3] the
 * unit Sun direction in each facet's frame (x along its first edge, z the
 * normal; may be NULL), solar = 1367 / r^2 (may be NULL).                   */
int btpm_geometry_illumination(btpm_ctx *ctx, const btpm_geometry *g, int *shadow, float *illum,
                               float *sun_facet, float *solar);

/* ---- view factors ------------------------------------------------------------
 * Facet view factors of a shape: replaces the Shape_Selfheating executable
 * (3905/Shape_Selfheating/main.cpp: initialise :750, selfheating :628, hidden
 * :187, betterviewfactors :374, the totals of main :817-835), whose outputs
 * (_selfheating_list.txt, _selfheating_vector_list.txt, _photovectors.txt,
 * :837-888) the Thermal stage reads.  Float arithmetic as the reference.    */
typedef struct {
    int nv, nf;
    const float *vtx;      /* [nv][3] vertex positions (the producer reads floats) */
    const int *facets;     /* [nf][3] zero-based vertex indices                     */
    int nsub;              /* sub-facet split per edge of refined pairs (reference
                              n = 10, main.cpp:15); 0 = no refinement, at most 32 */
    int cull;              /* 1: drop pairs hidden by a third facet (hidden())
                              as the reference; 0: keep every facing pair         */
    float refine_min;      /* refine pairs whose centroid view factor is at least
                              this (reference 0.05f, main.cpp:443)                */
} btpm_vf_mesh;
typedef struct btpm_vf btpm_vf;   /* the lists, resident on the device */
/* Builds the lists: row f holds, in increasing facet order, the facets that
 * f sees (list entry: facet, view factor).                                  */
int btpm_vf_build(btpm_ctx *ctx, const btpm_vf_mesh *mesh, btpm_vf **out);
/* facing pairs before the cull, list entries, refined entries              */
int btpm_vf_counts(const btpm_vf *vf, long long *n_facing, long long *nnz, long long *n_refined);
/* Copies the lists to host memory (any pointer may be NULL): rowptr [nf+1],
 * col [nnz], vf [nnz], dir [nnz][3] = the _selfheating_vector_list.txt value
 * vec / (VF A_f), vec [nnz][3] = the area-weighted VF vector, total [nf] =
 * the total view factor, photon [nf][3] = _photovectors.txt.               */
int btpm_vf_fetch(const btpm_vf *vf, int *rowptr, int *col, float *viewfactor, float *dir,
                  float *vec, float *total, float *photon);
void btpm_vf_free(btpm_vf *vf);

/* Shape->crater view factors: replaces calculate_surface_global_viewfactors
 * (3905/Thermal/NominalPole/main.cpp:1054-1181, 175706/Run-04/thermal.cpp:
 * 1477-1606) that the Thermal stage runs once per geometry before
 * load_GPU_roughness_globalillumination / _globalheating.                  */
typedef struct {
    int nv, nf;
    const double *vtx;     /* [nv][3] shape vertices (the stage reads doubles)    */
    const int *facets;     /* [nf][3]                                             */
    const int *rowptr;     /* [nf+1] facet view-factor rows                        */
    const float *vf;       /* [nnz] view factors                                   */
    const float *dir;      /* [nnz][3] vector-list directions                      */
    int ncv, ncf;                /* ncf at most 128                                 */
    const float *crater_vtx;     /* [ncv][3] crater model vertices                */
    const int *crater_facets;    /* [ncf][3]                                      */
    int m;                 /* rim-visibility samples per crater facet edge (m*m
                              samples; reference m = 10), at most 64             */
    float crater_radius;   /* crater rim radius in model units (reference 10)     */
} btpm_drf_in;
/* drf [nnz][ncf] (host memory): per list entry and crater sub-facet j,
 * VF vis_j cos_j / dir_z capped at 1.                                      */
int btpm_crater_viewfactors(btpm_ctx *ctx, const btpm_drf_in *in, float *drf);

/* Crater sunlight: replaces the crater part of surfacegeometry
 * (3905/Thermal/NominalPole/main.cpp:863-886) and the sfcpartialshadow kernel
 * (GPU_code.cu:152-210): per facet and step the cosine of the Sun on each
 * crater sub-facet and its sunlit fraction from m*m rim samples.  Outputs are
 * btpm_crater.cosang and .shadow.                                           */
typedef struct {
    int nf, ntau;
    const int *shadow;           /* [nf][ntau] facet lit flags (1 lit, 0 dark)      */
    const float *sun;            /* [nf][ntau][3] Sun vector in the facet frame, the
                                    reference's sfcillumvector (geometry producer
                                    with sun_scale 10000)                          */
    int ncv, ncf;                /* ncf at most 128                                 */
    const float *crater_vtx;
    const int *crater_facets;
    int m;                       /* rim samples per crater facet edge (10), <= 64   */
    float crater_radius;         /* 10                                              */
} btpm_crater_sun_in;
/* shadow, cosang: [nf][ntau][ncf] (host memory)                             */
int btpm_crater_sunlight(btpm_ctx *ctx, const btpm_crater_sun_in *in, float *shadow,
                         float *cosang);

/* ---- output files at scale ---------------------------------------------------
 * The stage's writers (T3905/main.cpp:2290-2341; 175706/Run-04/thermal.cpp:
 * 104-244, 3291-3381).  They return BTPM_ERR_ARG on an I/O failure.
 * Text temperatures, writeglobaltemperature's layout: vertex lines, facet
 * lines, then nf*ntau values one per line, facet-major, as the default
 * ostream prints them (%g); formatted on `threads` threads (0: all cores). */
int btpm_write_temperature_txt(const char *path, const double *vtx, int nv, const int *facets,
                               int nf, const float *stemp, int ntau, int threads);
/* binary_writesurfacetemperature: float[ncv][3], int[ncf][3], then
 * float[nf][ncf][ntau], from host temps [nf][ntau][ncf]                     */
int btpm_write_crater_dat(const char *path, const float *cvtx, int ncv, const int *cfv, int ncf,
                          const float *temps, int nf, int ntau);
/* the same file straight from a finished solve's device memory, streamed in
 * facet chunks through pinned buffers (no [nf][ntau][100] host copy)         */
int btpm_solve_write_crater_dat(btpm_solve *s, const char *path, const float *cvtx, int ncv,
                                const int *cfv, int ncf);
/* stitch_temperatures: part files (each a header plus whole facets of
 * ncf*ntau floats, consecutive facet ranges in order) joined into one file
 * with the last part's header, streamed                                      */
int btpm_stitch_crater_parts(const char *out_path, const char *const *parts, int nparts, int ncv,
                             int ncf, int ntau);

/* ---- observed fluxes ---------------------------------------------------------
 * The Observe stage's flux synthesis (175706/Run-04/observe_generator.cpp:
 * asteroidgeometry :530-698, surfacegeometry :465-529, observe :699-836),
 * downstream of the Thermal stage: per time step the flux of the smooth and
 * of the rough model at the observer in each band, after the colour
 * correction of the best-fitting colour temperature.  FP64.                 */
typedef struct {
    int nv, nf, ntau, nv_primary;
    const double *equpos;        /* [nv][3] each body's frame (shape units)          */
    const int *facets;           /* [nf][3]                                          */
    double rev_primary, rev_secondary;   /* revolutions per cycle                   */
    double hel[3];               /* system from the Sun: r (AU), ecl. lat, lon (rad) */
    double obs[3];               /* system from the observer: r (AU), lat, lon       */
    double rot[2];               /* pole latitude, longitude (rad)                   */
    double au_m;                 /* astronomical unit in metres (reference 149.6e9)  */
    const int *visible;          /* [nf][ntau] facet visibility (Visibility stage)   */
    const float *stemp;          /* [nf][ntau] smooth surface temperatures           */
    int rough;                   /* 1: crater (rough) fluxes as well                 */
    int ncv, ncf;
    const float *crater_vtx;     /* [ncv][3]                                         */
    const int *crater_facets;    /* [ncf][3]                                         */
    const float *crater_visible; /* [nf][ntau][ncf] sub-facet visible fractions      */
    const float *crater_temp;    /* [nf][ntau][ncf] crater temperatures              */
    const double *wavelength_um; /* [3] band wavelengths (reference WISE W2 W3 W4)   */
    const double *colour;        /* [3][tcount] colour corrections                   */
    int tmin, tcount;            /* colour temperatures tmin .. tmin+tcount-1 K      */
    double eps;                  /* emissivity                                       */
} btpm_observe_in;
/* smooth, rough: [ntau][3] (host memory)                                    */
int btpm_observe(btpm_ctx *ctx, const btpm_observe_in *in, double *smooth, double *rough);
/* The same synthesis with the temperatures taken from a finished solve in
 * device memory (no host copy of its [nf][ntau][100] crater temperatures):
 * the solve's shape has in->nf facets and in->ntau steps, and with in->rough
 * its crater model the in->ncf = 100 sub-facets; in->stemp and
 * in->crater_temp are not read (may be null).  Surface and crater
 * temperatures enter floored at +0 K, as btpm_solve_surface_temps and
 * btpm_solve_crater_temps return them, so the fluxes equal btpm_observe's on
 * those readbacks.  With obs_facet ([nf][ntau][3] floats, the observer's
 * direction in each facet's frame) the crater visibility is computed on the
 * device too, as btpm_crater_sunlight's fractions with the observer as the
 * source, in->visible as the facet flags, m x m rim samples and
 * crater_radius; in->crater_visible is then not read (may be null).  With
 * obs_facet null, in->crater_visible is used and m, crater_radius are not.
 * (No counterpart in the reference, whose Observe stage reads the Thermal
 * and Visibility stages' files.)                                           */
int btpm_observe_solve(btpm_ctx *ctx, const btpm_observe_in *in, btpm_solve *solve,
                       const float *obs_facet, int m, float crater_radius, double *smooth,
                       double *rough);

#ifdef __cplusplus
}
#endif
#endif
