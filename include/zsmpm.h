/* zsmpm: MI355X (gfx950) MLS-MPM substep on zpc's sparse block grid, C ABI.
 *
 * The entry points fill zpc's empty `rocm` execution space
 * (include/zensim/types/Property.h:46,51,86; resource/Resource.cpp:156-158)
 * for the operators of one explicit MLS-MPM substep. Each takes a policy
 * (device + stream + scratch, the counterpart of
 * CudaExecutionPolicy, cuda/execution/ExecutionPolicy.cuh:362-396) and
 * non-owning descriptors of device buffers in zpc's container layouts:
 *   table     HashTable<i32,3,int>: keys[T] (vec<i32,3>), indices[T], status[T],
 *             _cnt, _activeKeys[T]            (container/HashTable.hpp:40-58,200-205)
 *   grid      Grids<f32,3,4>{m:1, v:3, rhs:3}: element (ch, cell) of block b at
 *             (b*7 + ch)*64 + cell, cell = (x<<4)|(y<<2)|z
 *                                             (container/TileVector.hpp:108,397;
 *                                              geometry/Structure.hpp:849-860)
 *   particles Particles<f32,3>: x vec3, v vec3, C vec9, F vec9 (column-major),
 *             m, logJp, one array per attribute (geometry/Structurefree.hpp:23-290)
 *   buckets   IndexBuckets counts[B+1], offsets[B+1], indices[N]
 *                                             (container/IndexBuckets.hpp:11-66)
 * Naming follows zpc's py_interop `op__backend` exports
 * (py_interop/cuda/ExecutionPolicy.cpp:8-35); results are the int32 codes of
 * execution/AsyncRuntimeAbi.hpp:28-34, extended. Nothing throws across the
 * ABI. Calls are asynchronous on the policy's stream; device-side faults
 * (full table, missing block, ...) are reported by zsmpm_policy_sync.
 */
#ifndef ZSMPM_H
#define ZSMPM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ZSMPM_ABI_MAJOR 1
#define ZSMPM_ABI_MINOR 0

enum zsmpm_result {
  ZSMPM_OK = 0,
  ZSMPM_ERROR = -1,
  ZSMPM_INCOMPATIBLE_MAJOR = -2,
  ZSMPM_INSUFFICIENT_SIZE = -3,
  ZSMPM_UNSUPPORTED_OPERATION = -4,
  ZSMPM_INVALID_ARGUMENT = -5,
  ZSMPM_DEVICE_ERROR = -6,
  ZSMPM_TABLE_FULL = -7,       /* a probe walked the whole table / more keys than slots */
  ZSMPM_KEY_RANGE = -8,        /* a block key component outside [-2^20, 2^20) */
  ZSMPM_STENCIL_OUTSIDE = -9,  /* a particle stencil outside its bucket's 2x2x2 blocks */
  ZSMPM_MISSING_BLOCK = -10,   /* a needed block is not in the table (not activated/enlarged) */
  ZSMPM_CAPACITY = -11,        /* more blocks/particles than the buffers or reservation hold */
  ZSMPM_IO_ERROR = -12         /* a file could not be opened, read or written, or is not a valid checkpoint */
};

/* Versioned, size-prefixed struct header (execution/AsyncRuntimeAbi.hpp:12-27). */
typedef struct zsmpm_abi_header {
  uint32_t size;  /* sizeof the enclosing struct */
  uint16_t major; /* ZSMPM_ABI_MAJOR */
  uint16_t minor;
} zsmpm_abi_header;

typedef struct zsmpm_table {
  zsmpm_abi_header hdr;
  int32_t table_size; /* _tableSize, a power of two (evaluateTableSize, HashTable.hpp:88-91) */
  int32_t* keys;      /* [3*T] */
  int32_t* indices;   /* [T] */
  int32_t* status;    /* [T] */
  int32_t* cnt;       /* [1] */
  int32_t* active_keys; /* [3*T] */
} zsmpm_table;

typedef struct zsmpm_particles {
  zsmpm_abi_header hdr;
  int64_t n;
  float* x;     /* [3n] */
  float* v;     /* [3n] */
  float* C;     /* [9n] column-major */
  float* F;     /* [9n] column-major */
  float* m;     /* [n] */
  float* logJp; /* [n], NULL for models without plasticity */
  int32_t* id;  /* [n] particle ids carried by reordering, or NULL */
} zsmpm_particles;

typedef struct zsmpm_grid {
  zsmpm_abi_header hdr;
  float dx;
  int32_t capacity_blocks; /* blocks allocated in data */
  float* data;             /* [capacity_blocks * 7 * 64] */
} zsmpm_grid;

typedef struct zsmpm_buckets {
  zsmpm_abi_header hdr;
  int32_t capacity_blocks; /* counts/offsets hold capacity_blocks + 1 entries */
  int32_t* counts;
  int32_t* offsets;
  int32_t* indices; /* [n] */
} zsmpm_buckets;

enum zsmpm_model_kind { ZSMPM_FIXED_COROTATED = 0, ZSMPM_DRUCKER_PRAGER = 1 };

/* FixedCorotatedConfig / DruckerPragerConfig (physics/ConstitutiveModel.hpp:736-754). */
typedef struct zsmpm_model {
  zsmpm_abi_header hdr;
  int32_t kind;
  float volume;
  float E, nu;
  float cohesion, beta, yield_surface;
  int32_t volume_correction;
  float svd_fp64_gap; /* >0: FP64 SVD when min sigma gap < gap * sigma_1 */
} zsmpm_model;

enum zsmpm_collider_type { ZSMPM_STICKY = 0, ZSMPM_SLIP = 1, ZSMPM_SEPARATE = 2 };

/* Collider<AnalyticLevelSet<Plane>> with an identity transform (geometry/Collider.h:80-107). */
typedef struct zsmpm_plane_collider {
  int32_t type;
  float origin[3];
  float normal[3];
} zsmpm_plane_collider;

#define ZSMPM_MAX_COLLIDERS 8

enum zsmpm_geometry {
  ZSMPM_GEOMETRY_PLANE = 0,
  ZSMPM_GEOMETRY_CUBOID = 1,
  ZSMPM_GEOMETRY_SPHERE = 2,
  ZSMPM_GEOMETRY_CYLINDER = 3
};

/* Collider<AnalyticLevelSet<geometry>> with its rigid motion
 * (geometry/Collider.h:10-190; geometry/AnalyticLevelSet.h:11-225): the level
 * set's material point X sits at x = R s X + b, and the object's velocity at
 * x is omega x (x - b) + (dsdt / s)(x - b) + dbdt. A node with phi(X) < 0
 * takes that velocity (sticky) or loses the normal part of its velocity
 * relative to it (slip; separate only when approaching), the normal being
 * R n(X) (Collider::resolveCollision, Collider.h:81-107). Geometry in
 * material space: plane p0 = origin, p1 = normal (used as given); cuboid
 * p0 = min, p1 = max corner; sphere p0 = center, p1[0] = radius; cylinder
 * p0 = bottom, p1[0] = radius, p1[1] = length along `axis` (0..2). */
typedef struct zsmpm_collider {
  int32_t type;     /* zsmpm_collider_type */
  int32_t geometry; /* zsmpm_geometry */
  int32_t axis;     /* cylinder axis */
  float p0[3], p1[3];
  float R[9];       /* rotation, column-major */
  float omega[3];   /* angular velocity */
  float s, dsdt;    /* scale (> 0) and its rate */
  float b[3], dbdt[3]; /* translation and its rate */
} zsmpm_collider;

typedef struct zsmpm_policy_s* zsmpm_policy;

/* ABI version of the loaded library. */
int32_t zsmpm_abi_version(zsmpm_abi_header* out);

/* Policy for device `devid` on `stream` (NULL: the null stream)
 * (CudaExecutionPolicy, cuda/execution/ExecutionPolicy.cuh:362-396;
 *  py_interop/cuda/ExecutionPolicy.cpp:8-12 policy__device). */
int32_t zsmpm_policy__rocm(int32_t devid, void* stream, zsmpm_policy* out);
/* py_interop/cuda/ExecutionPolicy.cpp:13-15 del_policy__device. */
int32_t zsmpm_del_policy__rocm(zsmpm_policy pol);
/* Preallocate the policy's scratch for up to max_particles and a table of
 * table_size slots (get_temporary_memory_source, omp/execution/ExecutionPolicy.hpp:26). */
int32_t zsmpm_policy_reserve__rocm(zsmpm_policy pol, int64_t max_particles, int32_t table_size);
/* Wait for the stream; returns the first device-side error raised since the
 * last sync (cuda/execution/ExecutionPolicy.cuh:441-451 checks). */
int32_t zsmpm_policy_sync(zsmpm_policy pol);
/* Diagnostic counters after a sync: [0] new keys of the last insertion,
 * [2] particle runs of the last bucket sort, [3] FP64 SVDs, [4] closure fallbacks,
 * [5] max |v|^2 bits of the last grid update, [9] waves of 64 particles that
 * evaluated the closure network (all passes), [10] placement rounds of the
 * last zsmpm_bht_insert__rocm. out must hold 16 entries. */
int32_t zsmpm_policy_counters(zsmpm_policy pol, uint32_t* out);

/* Stages timed when profiling is on. */
enum zsmpm_stage {
  ZSMPM_STAGE_TABLE_CLEAR = 0,
  ZSMPM_STAGE_ACTIVATE = 1,
  ZSMPM_STAGE_ENLARGE = 2,
  ZSMPM_STAGE_BUCKET_SORT = 3,
  ZSMPM_STAGE_GRID_CLEAR = 4,
  ZSMPM_STAGE_P2G = 5,
  ZSMPM_STAGE_GRID_UPDATE = 6,
  ZSMPM_STAGE_G2P = 7,
  ZSMPM_STAGE_STRESS = 8, /* the per-particle constitutive part of P2G */
  ZSMPM_STAGE_COUNT = 9
};
/* policy.profile(bool) (execution/ExecutionPolicy.hpp:105-122; per-launch stream
 * timers cuda/Cuda.h:161-178): when enabled, every stage of the calls below is
 * bracketed by HIP events on the policy's stream. */
int32_t zsmpm_policy_profile(zsmpm_policy pol, int32_t enable);
/* Synchronises, then returns the accumulated milliseconds and call counts of
 * the first n stages (zsmpm_stage order); reset != 0 zeroes the totals. */
int32_t zsmpm_policy_stage_times(zsmpm_policy pol, double* ms, uint64_t* calls, int32_t n, int32_t reset);

/* CleanSparsity over range(tableSize), simulation/sparsity/SparsityOp.hpp:41-55. */
int32_t zsmpm_table_clear__rocm(zsmpm_policy pol, const zsmpm_table* table);
/* ComputeSparsity over range(n), SparsityOp.hpp:57-86: insert floor-div(floor(x/dx*... )
 * block keys; new keys are numbered in order of first occurrence (seq semantics). */
int32_t zsmpm_table_activate__rocm(zsmpm_policy pol, const zsmpm_table* table, const float* x, int64_t n,
                                   float dx, int32_t block_len, int32_t offset, float displacement);
/* zsmpm_table_activate__rocm without any allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Its scratch (dedup hash, new-key lists, first-occurrence bitmap, counters)
 * is one fixed array of the library's own device memory per GPU (about 46 MB,
 * held until the process ends), not policy scratch; the call makes no
 * hipMalloc, hipFree or allocating library call. Limits: table_size <= 2^21
 * slots and n <= 2^24 particles (ZSMPM_INSUFFICIENT_SIZE beyond). Results
 * equal zsmpm_table_activate__rocm's. Every call on one GPU shares that GPU's
 * array, whatever its policy: calls that overlap on one GPU (several streams
 * or host threads) race; calls ordered on one stream do not. Device-side
 * errors are reported by the next zsmpm_policy_sync of a policy that made
 * such a call. */
int32_t zsmpm_table_activate_noalloc__rocm(zsmpm_policy pol, const zsmpm_table* table, const float* x, int64_t n,
                                           float dx, int32_t block_len, int32_t offset, float displacement);
/* EnlargeSparsity over range(table.size()), SparsityOp.hpp:88-112. */
int32_t zsmpm_table_enlarge__rocm(zsmpm_policy pol, const zsmpm_table* table, const int32_t lo[3],
                                  const int32_t hi[3]);
/* HashTableView::query, container/HashTable.hpp:447-456 (modular probe wrap). */
int32_t zsmpm_table_query__rocm(zsmpm_policy pol, const zsmpm_table* table, const int32_t* keys, int64_t n,
                                int32_t* out);
/* index_buckets_for_particles: SpatiallyCount + exclusive_scan + SpatiallyDistribute,
 * simulation/particle/Query.tpp:30-56, SparsityOp.hpp:114-195; stable by particle index. */
int32_t zsmpm_bucket_sort__rocm(zsmpm_policy pol, const zsmpm_table* table, const float* x, int64_t n,
                                float dx, int32_t block_len, int32_t offset, float displacement,
                                const zsmpm_buckets* buckets);
/* CleanGridBlocks over Collapse{table.size(), 64}, simulation/grid/GridOp.hpp:54-69. */
int32_t zsmpm_grid_clear__rocm(zsmpm_policy pol, const zsmpm_grid* grid, const zsmpm_table* table);
/* P2GTransfer (APIC, quadratic) over range(n), simulation/transfer/P2G.hpp:45-127.
 * `buckets` is the particle bucket sort by (blockLen 4, offset -2, displacement 0.5)
 * keys; NULL computes it into policy scratch. Writes back logJp only. */
int32_t zsmpm_p2g__rocm(zsmpm_policy pol, const zsmpm_model* model, const zsmpm_particles* particles,
                        const zsmpm_table* table, const zsmpm_buckets* buckets, const zsmpm_grid* grid,
                        float dt);
/* ComputeGridBlockVelocity + ApplyBoundaryConditionOnGridBlocks per collider,
 * GridOp.hpp:71-164. fold_rhs != 0 first adds the rhs channels to momentum
 * (the explicit driver's convention). max_vel_sqr: device float or NULL;
 * receives max(*max_vel_sqr, max |v|^2 over nodes with mass) like zpc's
 * atomic_max on maxVel (any initial value, negative ones included). */
int32_t zsmpm_grid_update__rocm(zsmpm_policy pol, const zsmpm_grid* grid, const zsmpm_table* table, float dt,
                                const float extf[3], const zsmpm_plane_collider* colliders, int32_t ncolliders,
                                int32_t fold_rhs, float* max_vel_sqr);
/* GridAngularMomentum over Collapse{table.size(), 64}, simulation/grid/GridOp.hpp:216-259:
 * for every node whose channel m_channel is nonzero, with x = (blockkey * 4 +
 * coord(cell)) * dx, adds (double)(x cross mv) to sum[0..2] and (double)mv to
 * sum[3..5], mv = channels mv_channel..mv_channel+2. sum: 6 device doubles,
 * accumulated into (caller-initialised, like zpc's sumAngularMomentum). */
int32_t zsmpm_grid_angular_momentum__rocm(zsmpm_policy pol, const zsmpm_grid* grid, const zsmpm_table* table,
                                          int32_t m_channel, int32_t mv_channel, double* sum);
/* G2PTransfer (APIC) over range(n), simulation/transfer/G2P.hpp:43-84. */
int32_t zsmpm_g2p__rocm(zsmpm_policy pol, const zsmpm_particles* particles, const zsmpm_table* table,
                        const zsmpm_buckets* buckets, const zsmpm_grid* grid, float dt);
/* P2C2GTransfer (simulation/transfer/P2C2G.hpp:53-181), fixed-corotated and
 * Drucker-Prager: the particle -> cell -> node transfer. Each cell gathers the
 * particles within dx of its centre on every axis with the tent weight, C and
 * vol P F^T scaled per column by Dinv = 2 / (dx^2 - 2 Xrel^2) (Xrel: the
 * offset from the nearest node), the stress by -dt, summed with m C; each
 * cell then adds its moments to its 8 corner nodes with weight 1/8. Writes m
 * and mv (channels 0..3) into the cleared grid; the force channels are not
 * touched. The reference's cell-level IndexBuckets (index_buckets_for_particles,
 * simulation/particle/Query.tpp:9-56) are not needed: `buckets` is the same
 * block sort as for zsmpm_p2g__rocm (NULL computes it), and each particle's
 * stencil is scattered once. Drucker-Prager's logJp is written back once
 * (the reference writes it from every cell that gathers the particle). */
int32_t zsmpm_p2c2g__rocm(zsmpm_policy pol, const zsmpm_model* model, const zsmpm_particles* particles,
                          const zsmpm_table* table, const zsmpm_buckets* buckets, const zsmpm_grid* grid,
                          float dt);
/* PreG2C2PTransfer + G2C2PTransfer + PostG2C2PTransfer, G2C2P.hpp:58-271: v and
 * B = sum W v_i (x_i - x_p)^T gathered through the cells from the node
 * velocities (channels 1..3), B stored in C, F = (I + dt B Dinv) F at the old
 * position, x += v dt. */
int32_t zsmpm_g2c2p__rocm(zsmpm_policy pol, const zsmpm_particles* particles, const zsmpm_table* table,
                          const zsmpm_buckets* buckets, const zsmpm_grid* grid, float dt);
/* One explicit substep in the reconstructed driver's stage order: table clear,
 * activation (blockLen 4, offset -2, displacement 0.5), enlargement {0,0,0}..{2,2,2},
 * bucket sort, grid clear, P2G, rhs fold + grid velocity + colliders, G2P,
 * on the structures MPMSimulator's builder sets up (grid channels m, v, rhs;
 * the HashTable partition; maxVel; simulation/mpm/Simulator.cpp:44-130).
 * table, buckets and grid hold that substep's structures on return. */
int32_t zsmpm_substep__rocm(zsmpm_policy pol, const zsmpm_model* model, const zsmpm_particles* particles,
                            const zsmpm_table* table, const zsmpm_buckets* buckets, const zsmpm_grid* grid,
                            float dt, const float extf[3], const zsmpm_plane_collider* colliders,
                            int32_t ncolliders);

/* Vector::reorder(pol, buckets.indices, false_c) on every attribute
 * (container/Vector.hpp:617-646, gather form): out.attr[i] = in.attr[indices[i]],
 * out->id[i] = in->id ? in->id[indices[i]] : indices[i] when out->id is set.
 * in and out must not overlap; both hold in->n particles. */
int32_t zsmpm_reorder__rocm(zsmpm_policy pol, const zsmpm_buckets* buckets, const zsmpm_particles* in,
                            const zsmpm_particles* out);
/* zsmpm_substep__rocm followed by the driver-side reorder of the particles by
 * the substep's own bucket sort (Vector::reorder, container/Vector.hpp:617-646),
 * fused into G2P: the updated particles are written to `out` in bucket order,
 * so the next substep reads them block by block. P2G updates in->logJp in
 * place; the rest of `in` is read only. in and out must not overlap. */
int32_t zsmpm_substep_sorted__rocm(zsmpm_policy pol, const zsmpm_model* model, const zsmpm_particles* in,
                                   const zsmpm_particles* out, const zsmpm_table* table,
                                   const zsmpm_buckets* buckets, const zsmpm_grid* grid, float dt,
                                   const float extf[3], const zsmpm_plane_collider* colliders,
                                   int32_t ncolliders);

/* zsmpm_substeps_sorted__rocm flags */
#define ZSMPM_SUBSTEPS_FUSE_STRESS 1 /* G2P computes the next substep's stress (see below) */
#define ZSMPM_SUBSTEPS_GRAPH 2        /* replay a captured hipGraph of two substeps (see below) */
#define ZSMPM_SUBSTEPS_AUTOTUNE 4     /* pick the P2G / G2P workgroup counts on the first substeps */

/* nsteps substeps of zsmpm_substep_sorted__rocm, ping-ponging a -> b -> a ...:
 * the result is in b when nsteps is odd (*final_in_b = nsteps & 1); the
 * particle state after every substep is what the single-substep calls
 * produce. With ZSMPM_SUBSTEPS_FUSE_STRESS, G2P also computes the next
 * substep's stress from the F it stores (P2G.hpp:57-100), so the separate
 * per-particle constitutive pass runs once per call. a and b must not overlap.
 * With ZSMPM_SUBSTEPS_GRAPH (not combinable with FUSE_STRESS), the pair of
 * substeps a -> b -> a is captured once into a hipGraph (zpc launches every
 * kernel separately; graph replay has no counterpart there) and replayed
 * nsteps / 2 times on the policy's stream; an odd
 * remainder runs as one ordinary substep. The graph is kept on the policy and
 * captured again when any argument or the policy's scratch changes. Stage
 * profiling does not see graph replays.
 * With ZSMPM_SUBSTEPS_AUTOTUNE (and nsteps >= 8, not with FUSE_STRESS) the
 * first 8 substeps, which are the caller's own, run in pairs with four P2G
 * and four G2P launch configurations (tile and workgroup count) timed per
 * kernel; the policy keeps the fastest of each kernel for all later calls
 * (zsmpm_policy_launch_widths__rocm, zsmpm_policy_launch_tiles__rocm). */
int32_t zsmpm_substeps_sorted__rocm(zsmpm_policy pol, const zsmpm_model* model, const zsmpm_particles* a,
                                    const zsmpm_particles* b, const zsmpm_table* table,
                                    const zsmpm_buckets* buckets, const zsmpm_grid* grid, float dt,
                                    const float extf[3], const zsmpm_plane_collider* colliders, int32_t ncolliders,
                                    int32_t nsteps, int32_t flags, int32_t* final_in_b);

/* The general colliders of every later grid update on this policy (the
 * standalone update and the substep drivers, batches included, in each
 * scene's own coordinates), applied after the call's plane colliders in list
 * order (ApplyBoundaryConditionOnGridBlocks, GridOp.hpp:112-164). Copied;
 * a collider's position is the one given until the next call, which a driver
 * makes between substeps to move it. NULL or n = 0 removes them; at most
 * ZSMPM_MAX_COLLIDERS. */
int32_t zsmpm_policy_colliders__rocm(zsmpm_policy pol, const zsmpm_collider* colliders, int32_t n);

/* The P2G and G2P workgroup counts the policy launches with (autotuned or the
 * defaults). No counterpart in the reference (its launch shapes are fixed). */
int32_t zsmpm_policy_launch_widths__rocm(zsmpm_policy pol, int32_t* p2g_groups, int32_t* g2p_groups);

/* Sets the P2G and G2P workgroup counts of later launches, as the autotuner
 * would (0 restores the default; at most 1 << 20), e.g. to replay a tuned
 * configuration in another process. No counterpart in the reference. */
int32_t zsmpm_policy_set_launch_widths__rocm(zsmpm_policy pol, int32_t p2g_groups, int32_t g2p_groups);
/* The transfers of every later substep on this policy (zsmpm_substep__rocm,
 * zsmpm_substep_sorted__rocm, zsmpm_substeps_sorted__rocm, graphs included):
 * ZSMPM_TRANSFER_APIC (P2GTransfer / G2PTransfer, the default) or
 * ZSMPM_TRANSFER_CELL (P2C2GTransfer / G2C2PTransfer as zsmpm_p2c2g__rocm and
 * zsmpm_g2c2p__rocm compute them: the stored C is G2C2P's B). The cell
 * scheme runs one launch tile, and is not available for batches or with
 * ZSMPM_SUBSTEPS_FUSE_STRESS (ZSMPM_UNSUPPORTED_OPERATION). No counterpart in
 * zpc, whose driver calls the transfer functors it was built with. */
#define ZSMPM_TRANSFER_APIC 0
#define ZSMPM_TRANSFER_CELL 1
int32_t zsmpm_policy_transfer_scheme__rocm(zsmpm_policy pol, int32_t scheme);

/* The launch tiles of the sorted substep's P2G and G2P (no counterpart in zpc,
 * whose launches have one shape): P2G's chunk (576 particles over four waves,
 * three workgroups per CU, or 256 over two waves, six per CU) and G2P's
 * threads per bucket (128 or 256). ZSMPM_SUBSTEPS_AUTOTUNE picks them with the
 * workgroup counts; set 0 for the default (576, 128). Results do not depend
 * on them beyond float summation order. Other values: ZSMPM_INVALID_ARGUMENT. */
int32_t zsmpm_policy_launch_tiles__rocm(zsmpm_policy pol, int32_t* p2g_chunk, int32_t* g2p_threads);
int32_t zsmpm_policy_set_launch_tiles__rocm(zsmpm_policy pol, int32_t p2g_chunk, int32_t g2p_threads);

/* ------------------------------------------------------------------------
 * Batches of independent scenes (an addition: zpc advances one scene per
 * driver; no reference counterpart). The scenes share one set of particle
 * arrays, one table, one bucket sort and one grid, and every stage of a
 * substep is one launch over all of them. Scene s owns the particles
 * [first, first + count) (the scenes in order, first = the sum of the
 * earlier counts); its block keys are shifted by s * 1024 along axis 0, so
 * no block is shared and each block's scene follows from its key. A scene's
 * own particle block keys must lie in [-512, 511) along axis 0 (ZSMPM_KEY_RANGE
 * otherwise). Every scene has its own model, dx, dt, gravity and plane
 * colliders; a scene's particles, grid values and substep results are those
 * of zsmpm_substeps_sorted__rocm run on that scene alone (its tables list
 * the same keys, shifted, in the same relative order). The sorted driver
 * keeps each scene's particles inside its own range.
 * ---------------------------------------------------------------------- */
#define ZSMPM_MAX_SCENES 256

typedef struct zsmpm_batch_scene {
  zsmpm_abi_header hdr;
  int64_t first, count; /* particle range in the batch's arrays */
  zsmpm_model model;
  float dx, dt;
  float extf[3];
  int32_t ncolliders;
  zsmpm_plane_collider colliders[ZSMPM_MAX_COLLIDERS];
} zsmpm_batch_scene;

/* nsteps sorted substeps of every scene, ping-ponging a -> b -> a like
 * zsmpm_substeps_sorted__rocm (flags ZSMPM_SUBSTEPS_GRAPH and
 * ZSMPM_SUBSTEPS_AUTOTUNE as there; the graph holds two substeps of every
 * scene). a->id (when set) is carried by the reorder. logJp is required
 * when any scene is Drucker-Prager (other scenes' entries are carried).
 * grid->dx is not read. Counter 5 (max |v|^2) is the maximum over all
 * scenes. The learned closure is not applied to batches
 * (ZSMPM_UNSUPPORTED_OPERATION while one is selected). */
int32_t zsmpm_substeps_batched__rocm(zsmpm_policy pol, const zsmpm_batch_scene* scenes, int32_t nscenes,
                                     const zsmpm_particles* a, const zsmpm_particles* b, const zsmpm_table* table,
                                     const zsmpm_buckets* buckets, const zsmpm_grid* grid, int32_t nsteps,
                                     int32_t flags, int32_t* final_in_b);

/* ------------------------------------------------------------------------
 * Learned stress closure (an addition: zpc has no learned constitutive path).
 * For Drucker-Prager particles it replaces the SVD of the return mapping
 * (physics/ConstitutiveModel_Vol_dP.hpp:241-325, called from
 * simulation/transfer/P2G.hpp:84-100): with A = C - I, C = F^T F, n = |A|_F
 * and U = A / n, an MLP 9 -> 64 -> 64 -> 7 (GELU, bfloat16 operands on MFMA,
 * float accumulation) predicts E = log(C)/2 as n (psi0 I + psi1 U + psi2 U^2)
 * (and the projected strain and plastic multiplier, outputs 3..6); one Newton
 * step with series exp/log refines X = log C; the return mapping is applied to
 * E in tensor form and vol P F_proj^T = vol F S F^T, S = tau C^-1. Particles
 * whose a-posteriori residual bound exceeds tol (or with |D| >= 1/2,
 * |log C| > 1, det F <= 0) take the SVD path; their count accumulates in
 * counter 4 (zsmpm_policy_counters).
 *
 * Parameters (float32, ZSMPM_MLP_PARAMS): W1[9][64], b1[64], W2[64][64],
 * b2[64], W3[64][7], b3[7] (W[in][out]). Inputs per sample (9): tr U, tr U^3,
 * det U, log(n)/8 + 1, min(n, 1), clamp(logJp/n, -8, 8)/8, the friction
 * coefficient k = yield_surface (3 lambda + 2 mu) / (2 mu),
 * clamp(cohesion/n, -8, 8)/8, lambda / (lambda + 2 mu). Training records (10):
 * the eigenvalues m_i of U, e_i / n, lp_i / n (principal Hencky and projected
 * strains) and clamp(dg/n, -8, 8)/8. The loss per sample is
 * sum_i (psi0 + psi1 m_i + psi2 m_i^2 - e_i/n)^2 + the same for the projected
 * strain + (y6 - dg feature)^2; a pass minimises its mean over the samples.
 * ---------------------------------------------------------------------- */
#define ZSMPM_MLP_INPUTS 9
#define ZSMPM_MLP_HIDDEN 64
#define ZSMPM_MLP_OUTPUTS 7
#define ZSMPM_MLP_RECORD 10
#define ZSMPM_MLP_PARAMS 5255

#define ZSMPM_CLOSURE_ALWAYS_NETWORK 1 /* evaluate the network in every wave (see below) */

typedef struct zsmpm_closure {
  zsmpm_abi_header hdr;
  const float* params; /* [ZSMPM_MLP_PARAMS] network, or NULL: Newton from X0 = 0 */
  float tol;           /* relative residual bound above which the SVD path runs (1e-4) */
  int32_t flags;       /* ZSMPM_CLOSURE_*; by default a wave of 64 particles skips the
                          network when one step from X0 = 0 certifies all of them a priori
                          (|C - I|^3 / (4 (1 - |C - I|) (1 - |C - I|/2)) <= tol / 2) */
} zsmpm_closure;

/* Selects the closure of every later stress pass of Drucker-Prager models on
 * this policy (the plain SVD return mapping when closure is NULL). The weights
 * are read by each pass, not copied. Not combinable with
 * ZSMPM_SUBSTEPS_FUSE_STRESS (ZSMPM_UNSUPPORTED_OPERATION). */
int32_t zsmpm_policy_closure__rocm(zsmpm_policy pol, const zsmpm_closure* closure);

/* The stress pass on plain arrays: PF[9 i] = vol P F^T of F[9 i] (column-major)
 * with the policy's closure; logJp is updated (Drucker-Prager). The per-
 * particle part of P2G (P2G.hpp:57-100; compute_stress_* of
 * physics/ConstitutiveModel_Vol_dP.hpp:10-46, 241-325). */
int32_t zsmpm_stress__rocm(zsmpm_policy pol, const zsmpm_model* model, const float* F, float* logJp, int64_t n,
                           float* PF);

/* Training data from the plain path: features[9 i] and records[10 i] of
 * (F[9 i], logJp[i]), the targets from the device SVD return mapping. */
int32_t zsmpm_mlp_harvest__rocm(zsmpm_policy pol, const zsmpm_model* model, const float* F, const float* logJp,
                                int64_t n, float* features, float* records);

/* outputs[7 i] = MLP(features[9 i]). */
int32_t zsmpm_mlp_infer__rocm(zsmpm_policy pol, const float* params, const float* features, int64_t n,
                              float* outputs);

/* Gradient of the mean loss over the n samples (grad: ZSMPM_MLP_PARAMS floats)
 * and the loss (1 float, may be NULL), both device pointers. */
int32_t zsmpm_mlp_grad__rocm(zsmpm_policy pol, const float* params, const float* features, const float* records,
                             int64_t n, float* grad, float* loss);

/* One Adam step (beta 0.9 / 0.999, eps 1e-8; step counts from 1) on the batch:
 * params, adam_m and adam_v (ZSMPM_MLP_PARAMS floats each) are updated in
 * place; loss (device, may be NULL) receives the batch loss before the step. */
int32_t zsmpm_mlp_train_step__rocm(zsmpm_policy pol, float* params, float* adam_m, float* adam_v,
                                   const float* features, const float* records, int64_t n, float lr, int32_t step,
                                   float* loss);

/* ------------------------------------------------------------------------
 * Implicit step: G2P2GTransfer and ImplicitMPMSystem with the conjugate
 * gradient of math/linear/ConjugateGradient.hpp
 * (simulation/transfer/G2P2G.hpp:25-150; simulation/mpm/ImplicitMPM.hpp:13-159).
 * Grid vectors hold 3 floats per node, node = block * 64 + cell (zpc's dof
 * views over the grid); callers size them grid->capacity_blocks * 64 * 3; the
 * blocks the table holds are used. The grid's mass channel weights them.
 * ---------------------------------------------------------------------- */

/* G2P2GTransfer: per particle, C from the node velocities v, F = (I + dt C)
 * F_old (not stored), the model's vol P F^T times D = 4 / dx^2 (logJp read,
 * not written back), scattered into r: r[3 node + d] += W (PF[d] xixp0 +
 * PF[3 + d] xixp1 + PF[6 + d] xixp2). */
int32_t zsmpm_g2p2g__rocm(zsmpm_policy pol, const zsmpm_model* model, const zsmpm_particles* particles,
                          const zsmpm_table* table, const zsmpm_grid* grid, float dt, const float* v, float* r);
/* ImplicitMPMSystem::multiply: out = G2P2G(in) from zero, then out_i =
 * (out_i dt^2 + m_i) in_i on nodes with mass (ForceDtSqrPlusMass); other
 * nodes keep the force. in and out must differ. */
int32_t zsmpm_implicit_multiply__rocm(zsmpm_policy pol, const zsmpm_model* model, const zsmpm_particles* particles,
                                      const zsmpm_table* table, const zsmpm_grid* grid, float dt, const float* in,
                                      float* out);
/* ImplicitMPMSystem::project: with any boundary (the call's plane colliders,
 * the policy's general colliders and level-set boundary), nodes with mass
 * resolve their vector against each in turn and the others are cleared;
 * without one nothing changes. */
int32_t zsmpm_implicit_project__rocm(zsmpm_policy pol, const zsmpm_table* table, const zsmpm_grid* grid,
                                     const zsmpm_plane_collider* colliders, int32_t ncolliders, float* dof);
/* ImplicitMPMSystem::precondition: out_i = in_i / m_i on nodes with mass;
 * other entries of out are left alone. */
int32_t zsmpm_implicit_precondition__rocm(zsmpm_policy pol, const zsmpm_table* table, const zsmpm_grid* grid,
                                          const float* in, float* out);
/* ConjugateGradient::solve on that system: r = project(b - A x), q = M r,
 * p = q; while sqrt(r.q) > min(rel_tol sqrt(r0.q0), tol) and fewer than
 * max_iters iterations: temp = project(A p), alpha = r.q / temp.p,
 * x += alpha p, r -= alpha temp, q = M r, beta = r.q / previous, p = q +
 * beta p. zpc's defaults: tol 1e-6, rel_tol 0.5, max_iters 1000. x (device)
 * holds the start and receives the result; iters and residual (host, may be
 * NULL) the iterations taken and sqrt(r.q). Dot products are summed in
 * double and read back each iteration (the reference sums in float).
 * Synchronises the policy's stream. */
int32_t zsmpm_implicit_solve__rocm(zsmpm_policy pol, const zsmpm_model* model, const zsmpm_particles* particles,
                                   const zsmpm_table* table, const zsmpm_grid* grid, float dt,
                                   const zsmpm_plane_collider* colliders, int32_t ncolliders, float* x, const float* b,
                                   int32_t max_iters, float tol, float rel_tol, int32_t* iters, float* residual);

/* ------------------------------------------------------------------------
 * Bucketed hash table and SparseGrid
 * ---------------------------------------------------------------------- */

/* bht<i32, 3, int, 16> (container/Bht.hpp:16-310): buckets of 16 slots, three
 * universal hashes; a key goes to the first of its buckets (hf0, hf1, hf2)
 * that holds at most 14 keys, in the bucket's first empty slot. Storage keys
 * are vec<i32,3> padded to 16 bytes (x, y, z, pad); empty slots hold
 * 0x3f3f3f3f in every component. table_size is a multiple of 16
 * (zsmpm_bht_table_size); keys must be 16-byte aligned. hash holds (hashx,
 * hashy) of hf0, hf1, hf2 (zsmpm_bht_hash_params). Key components must lie
 * in [-2^20, 2^20) (ZSMPM_KEY_RANGE at the next sync otherwise). */
typedef struct zsmpm_bht {
  zsmpm_abi_header hdr;
  int32_t table_size;
  int32_t* keys;        /* [4 * table_size] */
  int32_t* indices;     /* [table_size] */
  int32_t* status;      /* [table_size] */
  int32_t* cnt;         /* [1] */
  int32_t* active_keys; /* [3 * table_size] (_activeKeys) */
  int32_t* success;     /* [1] (_buildSuccess) */
  uint32_t hash[6];
} zsmpm_bht;

#define ZSMPM_BHT_SEED 2    /* std::mt19937 rng(2) of bht's constructor (Bht.hpp:160-166) */
#define ZSMPM_BHT_SERIAL 1  /* zsmpm_bht_insert__rocm: one thread inserts the keys one by one */

/* evaluateTableSize (Bht.hpp:153-157): 2 next_pow2(n) rounded up to the next
 * multiple of 16 (by 16 when it is one); 0 for n = 0. Host function. */
int64_t zsmpm_bht_table_size(int64_t expected_entries);
/* The hash parameters bht draws from std::mt19937(seed) (Bht.hpp:160-166,
 * universal_hash(std::mt19937&), container/Bcht.hpp:39-43). Host function. */
void zsmpm_bht_hash_params(uint32_t seed, uint32_t out[6]);
/* bht::reset(clearCnt = true) (Bht.hpp:316-320): every key empty, status -1,
 * count 0, success 1; indices are kept. */
int32_t zsmpm_bht_clear__rocm(zsmpm_policy pol, const zsmpm_bht* table);
/* BHTView::insert for n keys (Bht.hpp:426-548), with the result of inserting
 * them one after another in input order: repeats and keys already present
 * are skipped, new keys take the next indices in order of first occurrence
 * and their slots are the host path's. out_index (device, may be NULL)
 * receives what insert returns for each input: the new index, -1 for a key
 * found, INT32_MIN where no bucket had room or the index comes within 20 of
 * the table size (both also raise ZSMPM_TABLE_FULL and clear success; the
 * table is then unspecified). flags ZSMPM_BHT_SERIAL: the sequential
 * insertion itself on one thread (the reference mode the default is checked
 * against). Synchronises the policy's stream. zpc's CUDA insert places keys
 * by atomics in an order that varies from run to run. */
int32_t zsmpm_bht_insert__rocm(zsmpm_policy pol, const zsmpm_bht* table, const int32_t* keys, int64_t n,
                               int32_t* out_index, int32_t flags);
/* BHTView::query / tile_query (Bht.hpp:551-620): out[i] = the index of keys[i]
 * or -1. */
int32_t zsmpm_bht_query__rocm(zsmpm_policy pol, const zsmpm_bht* table, const int32_t* keys, int64_t n,
                              int32_t* out);

/* SparseGrid<3, f32, 8> (geometry/SparseGrid.hpp:16-190): block origins
 * (multiples of 8) in the bht; cell values in a TileVector<f32, 512>: block b,
 * channel c, cell k at data[(b * num_channels + c) * 512 + k], k = (x * 8 +
 * y) * 8 + z; the index-to-world transform i2w (4 x 4, row-major, row-vector
 * convention: world = [X, 1] i2w, translation in row 3) and its inverse w2i;
 * the background value of inactive cells. */
typedef struct zsmpm_sparse_grid {
  zsmpm_abi_header hdr;
  zsmpm_bht table;
  float* data;             /* [capacity_blocks * num_channels * 512] */
  int32_t num_channels;
  int32_t capacity_blocks;
  float i2w[16];
  float w2i[16];
  float background;
} zsmpm_sparse_grid;

/* SparseGridView::insert(x) for n world positions (SparseGrid.hpp:439-446):
 * the block of floor(w2i x + 1/2), with zsmpm_bht_insert__rocm's semantics. */
int32_t zsmpm_spg_activate__rocm(zsmpm_policy pol, const zsmpm_sparse_grid* grid, const float* x, int64_t n);
/* wSample (width 1) / wPack (width 2 or 3) of channels [channel, channel +
 * width) at n world positions (SparseGrid.hpp:458-518, kernel_e::linear):
 * trilinear over the 2^3 cells around w2i x, the background where a block is
 * inactive. out: [width * n]. */
int32_t zsmpm_spg_sample__rocm(zsmpm_policy pol, const zsmpm_sparse_grid* grid, int32_t channel, int32_t width,
                               const float* x, int64_t n, float* out);
/* LevelSetBoundary<SparseLevelSet> (geometry/Collider.h:204-235) applied by
 * every later grid update of the policy after its plane and analytic
 * colliders (ApplyBoundaryConditionOnGridBlocks, GridOp.hpp:127-164): where
 * the grid's channel sdf_channel, sampled at the node, is negative, the node
 * takes the sampled velocity channels v_channel..+2 (zero for -1) when
 * sticky, or loses the normal part of its velocity relative to it (slip;
 * separate only when approaching), the normal from central differences of
 * the signed distance with eps = voxelSize(0) / 4 (SparseGrid.hpp:524-538).
 * The grid's buffers are read by every update, not copied. NULL removes it. */
int32_t zsmpm_policy_levelset_boundary__rocm(zsmpm_policy pol, const zsmpm_sparse_grid* grid, int32_t sdf_channel,
                                             int32_t v_channel, int32_t type);

/* ------------------------------------------------------------------------
 * Particle layout and IO adapters
 * ---------------------------------------------------------------------- */

/* Particles held in a TileVector<f32, L> (container/TileVector.hpp:15-1160):
 * element i, channel c at data[(i / L * num_channels + c) * L + i % L]
 * (TileVector.hpp:108,397,911-913). Each attribute is named by its first
 * channel (the property's tag offset, getPropertyOffset), -1 when the tile
 * set does not hold it; widths are x, v: 3, C, F: 9, m, logJp, id: 1 (id holds
 * int32 bits). The channel ranges of the named attributes must not overlap.
 * matrix_order ZSMPM_ROW_MAJOR: the 9 channels of C and F are a vec<f32,3,3>
 * as TileVector's pack<3,3> / tuple<9> read and write it (channel 3r + c is
 * entry (r, c)); ZSMPM_COLUMN_MAJOR: in the order of zsmpm_particles' rows. */
enum zsmpm_matrix_order { ZSMPM_COLUMN_MAJOR = 0, ZSMPM_ROW_MAJOR = 1 };

typedef struct zsmpm_tile_particles {
  zsmpm_abi_header hdr;
  int64_t n;            /* elements (TileVector::size()) */
  int32_t lane_width;   /* L: a power of two in [1, 1024]; zpc's particle tiles use 32 */
  int32_t num_channels; /* channels per element (TileVector::numChannels()) */
  float* data;          /* [ceil(n / L) * num_channels * L] */
  int32_t x, v, C, F, m, logJp, id; /* first channel of each attribute, or -1 */
  int32_t matrix_order; /* zsmpm_matrix_order of the C and F channels */
} zsmpm_tile_particles;

/* dst's arrays <- the attributes src holds (every attribute with a channel
 * >= 0 needs its dst array; dst arrays of absent attributes are left as they
 * are); dst->n == src->n. One pass over HBM. zpc has no such adapter: its
 * Particles keep one Vector per attribute (geometry/Structurefree.hpp:23-290)
 * and its TileVector particle sets (particleBins, :226) are filled attribute
 * by attribute through TileVector::pack / tuple. */
int32_t zsmpm_particles_from_tiles__rocm(zsmpm_policy pol, const zsmpm_tile_particles* src,
                                         const zsmpm_particles* dst);
/* The reverse: the named channels of dst <- src's arrays; other channels and
 * the lanes of the last tile past n are left as they are. */
int32_t zsmpm_particles_to_tiles__rocm(zsmpm_policy pol, const zsmpm_particles* src,
                                       const zsmpm_tile_particles* dst);

/* TileVector::reorderTiles (container/TileVector.hpp:644-680): whole tiles of
 * lane_width * num_channels floats; gather (scatter = 0): dst tile t = src
 * tile map[t]; scatter: dst tile map[t] = src tile t. map holds num_tiles
 * device int32 in [0, num_tiles) (ZSMPM_CAPACITY at the next sync
 * otherwise); src and dst must not overlap. */
int32_t zsmpm_tiles_reorder__rocm(zsmpm_policy pol, const float* src, float* dst, const int32_t* map,
                                  int64_t num_tiles, int32_t lane_width, int32_t num_channels, int32_t scatter);

/* Checkpoints: the particle arrays a descriptor names and, optionally, the
 * HashTable, in one binary file for resume. Layout (little-endian): a 48-byte
 * header (magic "ZSMPMCKP", u32 version 1, u32 field bits ZSMPM_CK_*, i64 n,
 * i32 table size or 0, i32 0, f64 time, i64 step); the particle arrays in the
 * order x, v, C, F, m, logJp, id (those present, raw); the table in zpc's
 * serialisation order (keys, indices, status, table size, cnt, active keys;
 * container/HashTable.hpp:576-590); a trailing u64 checksum (FNV-1a 64 over
 * the preceding bytes as little-endian 64-bit words, the tail bytes singly).
 * Save and load synchronise the policy's stream and return when the file is
 * written or the device arrays hold its contents. */
enum zsmpm_checkpoint_fields {
  ZSMPM_CK_X = 1, ZSMPM_CK_V = 2, ZSMPM_CK_C = 4, ZSMPM_CK_F = 8,
  ZSMPM_CK_M = 16, ZSMPM_CK_LOGJP = 32, ZSMPM_CK_ID = 64
};

typedef struct zsmpm_checkpoint_info {
  int64_t n;
  int32_t table_size; /* 0: no table in the file */
  int32_t fields;     /* ZSMPM_CK_* bits of the particle arrays in the file */
  double time;
  int64_t step;
} zsmpm_checkpoint_info;

/* x, v, C, F and m are required; logJp and id are saved when set; table may
 * be NULL. time and step are the caller's, returned by load. */
int32_t zsmpm_checkpoint_save__rocm(zsmpm_policy pol, const char* path, const zsmpm_particles* particles,
                                    const zsmpm_table* table, double time, int64_t step);
/* The header of a checkpoint, to size the buffers of a load. */
int32_t zsmpm_checkpoint_info__rocm(const char* path, zsmpm_checkpoint_info* out);
/* Fills the caller's arrays: particles must name exactly the file's arrays
 * with n equal to the file's, table (NULL when the file has none) the same
 * table size (ZSMPM_INVALID_ARGUMENT otherwise). A bad magic, version or
 * checksum, or a short file, is ZSMPM_IO_ERROR (the arrays may then hold part
 * of the file). out may be NULL. */
int32_t zsmpm_checkpoint_load__rocm(zsmpm_policy pol, const char* path, const zsmpm_particles* particles,
                                    const zsmpm_table* table, zsmpm_checkpoint_info* out);

/* write_partio / write_partio_with_stress / write_partio_with_grid
 * (io/ParticleIO.hpp:10-24, io/ParticleIO.cpp:7-76): a Houdini bgeo file of n
 * points at pos (device, 3 floats each) with, when attr is set, one point
 * attribute `attr_name` of attr_width (1: float, as "stress"; 3: vector, as
 * "force") floats per point (device). The layout is partio's bgeo writer's
 * (version 5, big-endian, position as a homogeneous 4-vector), the file
 * format zpc writes through partio. */
int32_t zsmpm_write_partio__rocm(zsmpm_policy pol, const char* path, const float* pos, int64_t n,
                                 const float* attr, int32_t attr_width, const char* attr_name);

#ifdef __cplusplus
}
#endif

#endif /* ZSMPM_H */
