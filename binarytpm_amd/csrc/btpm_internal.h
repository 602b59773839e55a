// btpm_internal.h -- device-visible descriptors shared by the HIP kernels and
// the C++ host runtime of libbtpm.
//
// HBM layout of one column set (a smooth body or the crater columns of one
// solve).  "Blocked SoA": time is cut into blocks of BTPM_BLK steps and each
// block stores the BTPM_BLK values of a column contiguously,
//     x[b][c][k]   (b = step / BLK, c = column, k = step % BLK),
// so a wavefront of 64 consecutive columns moves 64 x 32 B = 2 KB of
// contiguous memory per block (fully coalesced), while a single column's
// steps still come in 32-byte runs for the once-per-rotation gathers.
#pragma once
#include <stdint.h>

#define BTPM_NX 56          // depth nodes below the surface (reference NX)
#define BTPM_BLK 8          // steps per time block
#define BTPM_NS 100         // crater sub-facets per parent facet
#define BTPM_SMOOTH_WG 256  // columns per smooth-conduction workgroup
#define BTPM_CRATERS_WG 10  // craters per crater-conduction workgroup (1000 columns)
#define BTPM_CRATER_THREADS 1024
#define BTPM_PSTRIDE 64     // floats per column in the profile array (57 used)
// facet self-heating SpMM tiles: receivers per workgroup, emitters per
// staged chunk, steps per workgroup
#define BTPM_SH_ROWS 256
#define BTPM_SH_CHUNK 256
#define BTPM_SH_STEPS 64
#define BTPM_SH_STRIDE (BTPM_SH_STEPS + 4)  // LDS row stride of a staged chunk (CSR kernel)
#define BTPM_SH_WAVES (BTPM_SH_ROWS / 64)     // waves per self-heating workgroup
#define BTPM_SH_ROWS_T 256  // receivers per workgroup of the time-lane kernel (btpm_selfheat_rows)
#define BTPM_SH_STEPS_T 64  // steps per workgroup of the time-lane kernel (one per lane)
// the lane-interleaved (ELL) self-heating layout is used while its padded
// entry count stays within this multiple of the view-factor count
#define BTPM_SH_ELL_MAX_PAD 4
// halves per packed shape->crater view-factor tile: [hi, lo][112 sub-facets][32 entries]
#define BTPM_DRF_TILE (2 * 112 * 32)
// doubles per facet record of the shadow tests (geometry.hip)
#define BTPM_GEOM_REC 20
// floats per facet record of the view-factor producer (viewfactor.hip)
#define BTPM_VF_REC 28
// the facing test acosf(c) <= float(pi)/2 of the view-factor producer holds
// exactly for BTPM_VF_COS_MIN <= c <= 1 under glibc's acosf (checked over
// every float); NaN fails both
#define BTPM_VF_COS_MIN (-0x1.bbbd2cp-24f)
// largest sub-facet split per edge of a refined pair (n*n sub-facets)
#define BTPM_VF_NSUB_MAX 32
// the Observe stage's test acos(c) > pi/2 (double pi = 4 atan 1) holds
// exactly for -1 <= c <= BTPM_OBS_COS_BEHIND under glibc's acos (checked by
// bisection and on 2e6 doubles each side of the boundary)
#define BTPM_OBS_COS_BEHIND (-0x1.cb3b399d747f5p-55)
// largest accepted r = dtau/dz^2 (FTCS is stable for r <= 1/2): the
// conduction kernels integrate the depth profile in a variable scaled by
// (1 - 2r)^-k over 8-step blocks, which must stay inside the FP32 range
// ((1 - 2r)^-8 * 1e4 K < 3e38 needs 1 - 2r > 5.6e-5)
#define BTPM_R_MAX 0.49995

// kind of a column set
enum { BTPM_SET_SMOOTH = 0, BTPM_SET_CRATER = 1 };

// A set of columns advanced together rotation by rotation.
struct btpm_set_dev {
  int kind;           // BTPM_SET_SMOOTH or BTPM_SET_CRATER
  int ncols;          // columns (crater: ngroups * 100)
  int ngroups;        // freeze groups: smooth = ncols, crater = parent facets
  int ntau;           // steps per cycle
  int nb;             // time blocks = ceil(ntau / BLK)
  int max_rev;        // rotation cap
  int check_after;    // freeze test when rev > check_after
  int inst;           // 1: instantaneous equilibrium (TI = 0)
  float es;           // eps * sigma
  float esia;         // ir_abs * eps * sigma (self-heating emission)
  float tacc;         // freeze tolerance (K)
  float newton_tol;   // Newton stop tolerance (K)
  int newton_cap;     // Newton iteration cap
  float vf_unscale;   // crater: 2^-s undoing the power-of-two scale of vfA
  // arrays (device pointers)
  const float* src;   // [nb][ncols][BLK] absorbed flux (1-A)F (+RGH)
  float* T;           // [nb][ncols][BLK] surface T, last rotation (in place)
  float* prof;        // [ncols][PSTRIDE] depth profile, K (node i at +i)
  const float* r;     // [ngroups] FTCS number dtau/dz^2
  const float* m1;    // [ngroups] Gamma sqrt(2 pi n / P) / dz
  uint8_t* frozen;    // [ngroups]
  int* order;         // [ngroups] groups to integrate; the status kernel compacts
                      // the unfrozen ones to the front after every rotation
  float* dmax;        // [ngroups] max |dT| of the last executed rotation
  float* qext;        // [nb][ncols][BLK] total flux incl. facet self-heating (or null)
  float* emis;        // [ncols][nb*BLK] esia T^4 for the next rotation (or null),
                      // facet-major for the self-heating SpMM's coalesced reads
  const float* vfA;   // crater: fp16 hi/lo view-factor tiles (see crater kernel)
  // status, written by the kernels
  int* tile_active;   // [ntiles] active groups after this rotation's test
  int* tile_work;     // [ntiles] columns integrated during the rotation
  int ntiles;
  int tile_cols;      // columns per tile (smooth) / groups per tile (crater)
};

// per-set progress, device resident, advanced by btpm_status_kernel
struct btpm_status_dev {
  int rev;            // rotations completed
  int active;         // groups not frozen
  int done;           // 1 once every group froze or rev reached max_rev
  int converged;      // 1 if done by freezing (not by the cap)
  long long colsteps; // integrated column-steps (executed, excluding frozen)
  int norder;         // leading entries of the set's order[] to integrate
  int pad_;
};

// one workgroup's work item in a batched launch
struct btpm_tile {
  int set;            // index into the set table
  int tile;           // tile index inside the set
};
