/* vortegrid.h -- C-ABI of the MI355X-native velocity-from-vorticity path.
 *
 * Drop-in boundary for VorteGrid's VortonSim path (mijagourlay/VorteGrid,
 * Samples/MjgIntelFluidDemo_Part20/): direct Biot-Savart, the NestedGrid
 * influence treecode, the vector-Poisson multigrid solve with treecode
 * Dirichlet boundaries, and trilinear grid interpolation for tracer
 * advection.  Every entry point names the reference routine it replaces.
 *
 * Conventions
 *  - Plain pointers and sizes only.  Particle/grid buffers passed to vg_*
 *    compute calls are DEVICE pointers owned by the caller (allocate them
 *    with vg_malloc or any HIP allocator); the library never keeps them
 *    beyond the call except where noted (vg_sim_set_vortons mirrors
 *    VortonSim::SetVortons, vortonSim.h:329, and keeps the view).
 *  - Vec3 arrays are "strided float triples": element i starts at
 *    ptr[i * stride] (stride in floats).  stride 3 = packed Vec3 array;
 *    stride 18 = the 72-byte Particle record (particle.h:425-446) with
 *    pos at float offset 0, velocity at 3, angular velocity at 9, size at 13.
 *  - Work is enqueued on the given stream (vg_stream, a hipStream_t; NULL =
 *    the legacy default stream) and is asynchronous unless the function says
 *    it returns host data.
 *  - Status codes instead of the reference's debug ASSERTs: empty inputs are
 *    VG_OK with a zero field (vortonSim.cpp:2668-2675); zero-extent grids are
 *    a no-op (pclOpAdvect.cpp:138-142); out-of-domain positions are clamped
 *    to the boundary cell (the reference's behaviour there is undefined).
 */
#ifndef VORTEGRID_H
#define VORTEGRID_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum vg_status {
    VG_OK = 0,
    VG_EINVAL = 1,   /* bad argument (null pointer, bad size, bad enum) */
    VG_EMPTY = 2,    /* reserved: nothing to do is reported as VG_OK */
    VG_EDEVICE = 3,  /* HIP runtime error; see vg_last_error() */
    VG_ENOMEM = 4,   /* device allocation failed */
    VG_ESTATE = 5    /* call order violated (e.g. no tree built yet) */
} vg_status;

typedef void *vg_stream; /* hipStream_t */

/* Rsqrt policy of the pair kernels.  REFERENCE reproduces finvsqrtf
 * (Core/Math/math.h:206-216, one Newton step, <=1.75e-3 rel error) for
 * parity; ACCURATE uses the hardware reciprocal square root (~1 ulp). */
enum { VG_RSQRT_REFERENCE = 0, VG_RSQRT_ACCURATE = 1 };

/* Velocity-from-vorticity technique (vorton.h:37-47). */
enum {
    VG_TECH_DIRECT = 0,   /* VELOCITY_TECHNIQUE_DIRECT */
    VG_TECH_TREE = 1,     /* VELOCITY_TECHNIQUE_TREE */
    VG_TECH_POISSON = 2   /* VELOCITY_TECHNIQUE_POISSON_GAUSS_SEIDEL */
};

/* Grid geometry: UniformGridGeometry (uniformGrid.h:144-860).  Points are
 * x-fastest: offset = ix + np[0]*(iy + np[1]*iz) (uniformGrid.h:756-759). */
typedef struct vg_grid_desc {
    float min[3];     /* mMinCorner */
    float extent[3];  /* mGridExtent */
    float cell[3];    /* mCellExtent (spacing) */
    float cpe[3];     /* mCellsPerExtent */
    uint32_t np[3];   /* mNumPoints */
} vg_grid_desc;

/* Device view of a vorton array (Vorton = Particle, vorton.h:198-387). */
typedef struct vg_vortons {
    const float *pos;     /* strided positions */
    const float *angvel;  /* strided angular velocity = vorticity / 2 */
    const float *size;    /* strided diameters (mSize) or NULL */
    size_t stride;        /* floats between consecutive records */
    size_t n;
    float uniform_size;   /* diameter used when size == NULL */
} vg_vortons;

/* ------------------------------------------------------------ runtime */
const char *vg_last_error(void);
const char *vg_version(void);
vg_status vg_device_count(int *count);
vg_status vg_malloc(void **ptr, size_t bytes);
vg_status vg_free(void *ptr);
vg_status vg_memcpy_htod(void *dst, const void *src, size_t bytes);
vg_status vg_memcpy_dtoh(void *dst, const void *src, size_t bytes);
vg_status vg_memset(void *dst, int value, size_t bytes);
vg_status vg_stream_create(vg_stream *s);
vg_status vg_stream_destroy(vg_stream s);
vg_status vg_stream_synchronize(vg_stream s);
/* Kernel timing: when enabled, every kernel the library launches is
 * bracketed by HIP events on its own stream and its duration is tallied
 * by kernel name (read with vg_profile_get after a synchronize). */
vg_status vg_profile_enable(int on);
/* Times only the kernels whose names start with one of the comma-separated
 * prefixes (NULL or "": every kernel).  Each timed launch costs two event
 * records on its stream, so a timed frame selects the kernels it reports. */
vg_status vg_profile_select(const char *prefixes);
vg_status vg_profile_reset(void);
vg_status vg_profile_get(const char *kernel, double *total_ms, uint64_t *count);
/* Writes "name count total_ms\n" lines into buf (NUL-terminated). */
vg_status vg_profile_dump(char *buf, size_t cap);

/* ------------------------------------------------------------ geometry */
/* UniformGridGeometry::DefineShape (uniformGrid.h:259-322). */
vg_status vg_define_shape(vg_grid_desc *g, size_t num_elements,
                          const float vmin[3], const float vmax[3], int pow2);
/* UniformGridGeometry::Decimate (uniformGrid.h:361-376). */
vg_status vg_decimate(vg_grid_desc *dst, const vg_grid_desc *src, uint32_t dec);
/* NestedGrid::Initialize + PrecomputeDecimations (nestedGrid.h:69-85,
 * 312-391): writes up to max_layers layers and per-parent decimations
 * (dec[3*l + a]; dec of layer 0 is 0). */
vg_status vg_nested_layers(const vg_grid_desc *src, vg_grid_desc *layers,
                           uint32_t *dec, uint32_t max_layers, uint32_t *depth);
/* VortonSim::UpdateBoundingBox(min, max, bFinal=true) (vortonSim.cpp:
 * 1027-1096): pads the box by 3 vorton radii and calls DefineShape(n, ...,
 * pow2=true).  mn/mx are updated in place. */
vg_status vg_finalize_bounding_box(float mn[3], float mx[3], float vorton_radius,
                                   size_t num_vortons, vg_grid_desc *grid_template);
/* Auto-mollification (vortonSim.cpp:2458-2477). */
vg_status vg_auto_mollify(const vg_grid_desc *g, float vorton_radius,
                          float *spread, float *circ);

/* ------------------------------------------------------------ AABB */
/* PclOpFindBoundingBox::FindBoundingBox (pclOpFindBoundingBox.cpp:20-148):
 * grows mn/mx (host arrays, in/out) by the n strided positions.  Blocks
 * until the result is on the host. */
vg_status vg_find_bounding_box(const float *pos, size_t stride, size_t n,
                               float mn[3], float mx[3], vg_stream s);

/* ------------------------------------------------------------ direct sums */
/* VortonSim::ComputeVelocity_Direct (vortonSim.cpp:1144-1158) evaluated at
 * nq strided query points; out is packed xyz (3*nq floats).  FP32 with
 * tile-compensated (Kahan) accumulation. */
vg_status vg_direct_velocity(const vg_vortons *v, const float *q, size_t q_stride,
                             size_t nq, float spread, float circ, int rsqrt_policy,
                             float *out, vg_stream s);
/* No reference counterpart.  vg_direct_velocity for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc): it calls no allocator.  Its
 * transient buffers (the packed sources, 32 B per vorton, and the
 * source-split partials, 12 B per query and split) live in a fixed 32 MiB
 * device array of the library, one per GPU, held until the process ends.
 * Sizes that need more are refused with VG_EINVAL (1,048,576 vortons at
 * most, fewer when a small query set is split over the sources).  Every
 * call on a GPU uses that GPU's one array and overwrites all it reads, so
 * nothing needs clearing between calls, but calls that overlap on one GPU
 * (other streams, other threads) race on it: serialise them.  Results equal
 * vg_direct_velocity's bit for bit. */
vg_status vg_direct_velocity_noalloc(const vg_vortons *v, const float *q, size_t q_stride,
                                     size_t nq, float spread, float circ, int rsqrt_policy,
                                     float *out, vg_stream s);
/* VortonSim::ComputeVectorPotential_Direct (vortonSim.cpp:1113-1127). */
vg_status vg_direct_vecpot(const vg_vortons *v, const float *q, size_t q_stride,
                           size_t nq, float spread, float circ, int rsqrt_policy,
                           float *out, vg_stream s);
/* Direct velocity at every grid point (ComputeVelocityAtGridpoints_Slice,
 * DIRECT branch, vortonSim.cpp:2223-2293, with the 1-2eps spacing nudge). */
vg_status vg_direct_velocity_grid(const vg_vortons *v, const vg_grid_desc *g,
                                  float spread, float circ, int rsqrt_policy,
                                  float *out, vg_stream s);
/* FP64 accuracy reference on the GPU: same regularised law, exact
 * reciprocal square root, double accumulation.  out: 3*nq doubles. */
vg_status vg_direct_velocity_f64(const vg_vortons *v, const float *q,
                                 size_t q_stride, size_t nq, float spread,
                                 float circ, double *out, vg_stream s);

/* ------------------------------------------------------------ influence tree */
typedef struct vg_tree vg_tree;
vg_status vg_tree_create(vg_tree **t);
vg_status vg_tree_destroy(vg_tree *t);
/* VortonSim::CreateInfluenceTree (vortonSim.cpp:1605-1624) =
 * NestedGrid::Initialize + MakeBaseVortonGrid (:1192-1321) +
 * AggregateClusters (:1337-1569).  Cluster sums are accumulated in vorton
 * index order per cell, so layers match the reference bit for bit. */
vg_status vg_tree_build(vg_tree *t, const vg_grid_desc *grid_template,
                        const vg_vortons *v, vg_stream s);
vg_status vg_tree_depth(const vg_tree *t, uint32_t *depth);
/* Copies layer l to the host: grid descriptor, and (optional, may be NULL)
 * cluster position[3*cap], angular velocity[3*cap], size[cap].  Blocks. */
vg_status vg_tree_layer(const vg_tree *t, uint32_t l, vg_grid_desc *g,
                        float *pos, float *angvel, float *size);
/* First moments of layer l's clusters about their centroids (no reference
 * counterpart): dip[9*c + 3*b + e] = sum Omega_b (y - c)_e.  Blocks. */
vg_status vg_tree_layer_moments(const vg_tree *t, uint32_t l, float *dip);
/* ComputeVelocity_Tree from the root (vortonSim.cpp:1787-1905) at nq
 * strided query points. */
vg_status vg_tree_velocity(const vg_tree *t, const float *q, size_t q_stride,
                           size_t nq, float spread, float circ,
                           int rsqrt_policy, float *out, vg_stream s);
/* ComputeVelocityAtGridpoints_Slice, TREE branch (vortonSim.cpp:2223-2293):
 * every point of grid g (direct sum when the tree has one layer). */
vg_status vg_tree_velocity_grid(const vg_tree *t, const vg_grid_desc *g,
                                float spread, float circ, int rsqrt_policy,
                                float *out, vg_stream s);
/* Number of cluster interactions vg_tree_velocity_grid performs on grid g
 * (same walk, no pair arithmetic): the work measure for pair-rate and
 * FLOP-rate accounting.  Blocks until the count is on the host. */
vg_status vg_tree_interaction_count(const vg_tree *t, const vg_grid_desc *g,
                                    uint64_t *count, vg_stream s);
/* Accurate near-field treecode at every point of g (g must be the tree's
 * grid template).  No reference counterpart: replaces the TREE technique's
 * grid evaluation (vortonSim.cpp:2223-2293) when accuracy matters.  Grid
 * points are processed in 4x4x4 bricks; on every layer the cells within
 * near_k cells of the brick are opened, the other children of the parent's
 * near box are applied as clusters (order 0: monopole at the |w|-weighted
 * centroid, order 1: plus the first-moment correction), and the leaf near
 * box is summed over its vortons.  A tree whose grid leaves a cell without a
 * parent cluster (a cell count odd and above 3 on some layer, possible only
 * for grids that are not power-of-two shaped) is refused with VG_EINVAL:
 * the layers above never see that cell's vortons. */
vg_status vg_tree_velocity_grid_mac(const vg_tree *t, const vg_grid_desc *g, int near_k,
                                    int order, int expand_from, float spread, float circ,
                                    int rsqrt_policy, float *out, vg_stream s);
/* expand_from (>= 1; < 1 or > depth-2 disables it): clusters of layers >=
 * expand_from are applied through a local expansion about the brick centre
 * (monopole to second order, first moment to first order), one per lane and
 * summed over the wave, instead of at every point; 2 or 3 keeps the error of
 * the near-field treecode within ~10%. */
/* Work of vg_tree_velocity_grid_mac summed over all grid points: cluster
 * interactions and direct vorton pairs at points, and clusters expanded
 * (once per brick; may be NULL).  Blocks. */
vg_status vg_tree_mac_work(const vg_tree *t, const vg_grid_desc *g, int near_k, int expand_from,
                           uint64_t *clusters, uint64_t *pairs, uint64_t *expansions,
                           vg_stream s);
/* ComputeVectorPotential_Tree (vortonSim.cpp:1650-1768) at query points. */
vg_status vg_tree_vecpot(const vg_tree *t, const float *q, size_t q_stride,
                         size_t nq, float spread, float circ, int rsqrt_policy,
                         float *out, vg_stream s);
/* ComputeVectorPotentialAtGridpoints_Slice(boundariesOnly=true)
 * (vortonSim.cpp:2055-2138): writes the 6 faces of out (interior left
 * untouched).  copy_odd=1 reproduces the reference (odd x copies x-1);
 * copy_odd=0 computes every face point. */
vg_status vg_tree_vecpot_boundary(const vg_tree *t, const vg_grid_desc *g,
                                  float spread, float circ, int rsqrt_policy,
                                  int copy_odd, float *out, vg_stream s);
/* PopulateVorticityGridFromVortons(grid, scale) (vortonSim.cpp:2372-2436):
 * trilinear splat of vorticity*scale*V_vorton/V_cell onto the tree's leaf
 * grid geometry (= the grid template).  out: 3*capacity floats. */
vg_status vg_tree_splat_vorticity(const vg_tree *t, float vorton_radius,
                                  float scale, float *out, vg_stream s);
/* Particles::PopulateVelocityGrid (pclOpPopulateVelocityGrid.cpp:33-87):
 * velocities vel (strided, indexed like the vortons given to
 * vg_tree_build) averaged onto the leaf grid. */
vg_status vg_tree_populate_velocity_grid(const vg_tree *t, const float *vel,
                                         size_t vel_stride, float *out,
                                         vg_stream s);

/* ------------------------------------------------------------ grid operators */
/* UniformGrid::DownSample, FASTER_LESS_ACCURATE injection
 * (uniformGrid.h:1453-1512). */
vg_status vg_downsample_inject(const vg_grid_desc *lo, float *lo_field,
                               const vg_grid_desc *hi, const float *hi_field,
                               vg_stream s);
/* UniformGrid::UpSample(INTERIOR_ONLY) (uniformGrid.h:1637-1700): overwrites
 * interior fine points with trilinear samples of the coarse grid. */
vg_status vg_upsample_interior(const vg_grid_desc *hi, float *hi_field,
                               const vg_grid_desc *lo, const float *lo_field,
                               vg_stream s);
/* SolveVectorPoisson(soln, lap, steps, BC_DIRICHLET) (uniformGridMath.cpp:
 * 1369-1405): steps x (red zebra pass, black zebra pass) of SOR 1.72 with
 * x sequential in place (:1116-1188). */
vg_status vg_sor_solve(const vg_grid_desc *g, float *soln, const float *lap,
                       uint32_t steps, vg_stream s);
/* ComputeJacobian + ComputeCurlFromJacobian (uniformGridMath.cpp:207-451),
 * fused: no Mat33 grid is materialised. */
vg_status vg_curl(const vg_grid_desc *g, const float *a, float *curl, vg_stream s);
/* FP64 relative residual ||f - L u|| / ||f|| over interior points
 * (7-point Laplacian).  Blocks until the result is on the host. */
vg_status vg_residual_norm(const vg_grid_desc *g, const float *u, const float *f,
                           double *rel, vg_stream s);

/* ------------------------------------------------------------ multigrid */
typedef struct vg_multigrid vg_multigrid;
vg_status vg_mg_create(vg_multigrid **m);
vg_status vg_mg_destroy(vg_multigrid *m);
/* VortonSim::ComputeVectorPotential multigrid cascade (vortonSim.cpp:
 * 2523-2571): a0 holds the Dirichlet boundary on entry (its interior is
 * reset to 0, the cold start of :2531); negw0 = -vorticity.  Injection
 * restriction, replace-prolongation, steps_per_level SOR steps per visit. */
vg_status vg_mg_cascade(vg_multigrid *m, const vg_grid_desc *g, float *a0,
                        const float *negw0, uint32_t steps_per_level,
                        vg_stream s);
/* Residual-correction V-cycles (red-black Gauss-Seidel smoothing, full
 * weighting restriction, trilinear correction, FP64 coarsest solve and
 * FP64 residual norms) until ||f-Au||/||f|| <= tol or max_cycles.
 * Boundary values of a0 are kept (Dirichlet); the interior initial guess is
 * whatever a0 holds.  Reports the final relative residual and cycles. */
vg_status vg_mg_vcycle_solve(vg_multigrid *m, const vg_grid_desc *g, float *a0,
                             const float *negw0, double tol, uint32_t max_cycles,
                             double *rel_residual, uint32_t *cycles, vg_stream s);
/* The V-cycle solution is the compensated pair a0 + lo: an FP32 field alone
 * has a rounding floor of ~1e-4 in the relative residual at 129^3, so the
 * 1e-6 tolerance holds for the pair (each component summed in FP64).  This
 * returns the device pointer of lo (count = 3 * grid points floats), valid
 * until the next solve on m.  No reference counterpart (the reference's
 * cascade, vortonSim.cpp:2523-2571, never converges to a tolerance). */
vg_status vg_mg_vcycle_compensation(const vg_multigrid *m, const float **lo, size_t *count);
/* The cascade's pyramid (NestedGrid<Vec3>, vortonSim.h:338-343): layer l >= 1
 * of the vector potential (A) or of the negative vorticity (-w) as the last
 * vg_mg_cascade left it (device pointer, packed Vec3, valid until the next
 * cascade on m; layer 0 is the caller's a0 / negw0).  depth (may be NULL)
 * receives the number of layers that cascade filled: 1 + the coarse layers
 * it visited (those with min dimension > 2, vortonSim.cpp:2549-2561). */
enum { VG_MG_VECTOR_POTENTIAL = 0, VG_MG_NEGATIVE_VORTICITY = 1 };
vg_status vg_mg_layer(const vg_multigrid *m, int field, uint32_t layer, vg_grid_desc *g,
                      const float **data, uint32_t *depth);

/* ------------------------------------------------------------ particles */
/* AssignVelocityFromField (pclOpAdvect.cpp:132-152, slice :42-81): vel =
 * interp (gain == 1) or vel += gain*(interp - vel). */
vg_status vg_assign_velocity_from_field(const vg_grid_desc *g, const float *field,
                                        const float *pos, float *vel,
                                        size_t stride, size_t n, float gain,
                                        vg_stream s);
/* PclOpEvolve (pclOpEvolve.cpp:36-123): pos += vel * dt. */
vg_status vg_evolve(float *pos, const float *vel, size_t stride, size_t n,
                    float dt, vg_stream s);
/* Fused AssignVelocityFromField + Evolve for tracers: one read and one write
 * of each position.  vel may be NULL (velocity not stored; gain must then be
 * 1), otherwise it is read when gain != 1 and written. */
vg_status vg_advect_tracers(const vg_grid_desc *g, const float *field, float *pos,
                            float *vel, size_t stride, size_t n, float dt,
                            float gain, vg_stream s);

/* No reference counterpart: reorders n particle records in place so that
 * they follow the cells of grid g (x fastest, positions outside the grid
 * clamped to its boundary cells), keeping index order inside a cell (a
 * stable counting sort; deterministic).  A record is `stride` floats with
 * the position first (3 for packed positions, 18 for 72-byte Particle
 * records, which then move whole).  perm (device, n entries, may be NULL)
 * receives the permutation: new record j is old record perm[j], so that
 * other per-particle arrays can follow.  The reference keeps no particle
 * order (Particles::Kill swaps the last particle in, particle.cpp:103), so
 * an application may reorder its tracers; advected in cell order, their
 * grid gathers share cache lines (vg_advect_tracers on 4M tracers in
 * uniform random order gathers ~5 lines per tracer from the Infinity Cache). */
vg_status vg_sort_particles_by_cell(const vg_grid_desc *g, float *records, size_t stride,
                                   size_t n, uint32_t *perm, vg_stream s);

/* ------------------------------------------------------------ stretching and tilting */
/* ComputeJacobian (uniformGridMath.cpp:304-451): jac[9 o + 3 a + b] =
 * d v_b / d a (the Mat33 members x, y, z hold d/dx, d/dy, d/dz): central
 * 1/(2h) inside, one-sided 1/h on the faces, 1/h_z = 0 when
 * h_z <= FLT_EPSILON.  jac holds 9 * capacity floats. */
vg_status vg_jacobian(const vg_grid_desc *g, const float *vel, float *jac,
                      vg_stream s);
/* The per-vorton loop of VortonSim::StretchAndTiltVortons
 * (vortonSim.cpp:2796-2812): J = trilinear interpolation of the Jacobian
 * grid at the position (UniformGrid<Mat33>::Interpolate,
 * uniformGrid.h:1117-1148), s = Omega * J (transpose formulation,
 * mat33.h:131-145: s_a = J.a . Omega), Omega += 0.5 s dt.  A no-op when any
 * grid extent is zero (2-D domain, :2754-2759).  Positions and angular
 * velocities are strided separately (3 for packed arrays, 18 for 72-byte
 * Particle records).  order (device, may be NULL) is a permutation of
 * 0..n-1 giving the order in which vortons are processed: it changes no
 * value, only cache behaviour (the influence tree's cell order makes the
 * grid gathers of neighbouring threads hit the same cache lines). */
vg_status vg_stretch_tilt(const vg_grid_desc *g, const float *jac,
                          const float *pos, size_t pos_stride, float *angvel,
                          size_t ang_stride, size_t n, float dt,
                          const uint32_t *order, vg_stream s);
/* The same update from the velocity grid: the Jacobians of the 8 corners of
 * each vorton's cell are formed on the fly, so no Jacobian grid is written.
 * Equal to vg_jacobian + vg_stretch_tilt bit for bit. */
vg_status vg_stretch_tilt_from_velocity(const vg_grid_desc *g, const float *vel,
                                        const float *pos, size_t pos_stride,
                                        float *angvel, size_t ang_stride,
                                        size_t n, float dt,
                                        const uint32_t *order, vg_stream s);

/* ------------------------------------------------------------ viscous diffusion (PSE) */
/* PartitionVortons (vortonSim.cpp:3578-3607) / Particles::PartitionParticles
 * (particle.cpp:398-435): the tree's leaf cell lists are the reference's
 * ugVortonIndices on the grid template.  Vortons of cell c (indexed by its
 * minimum-corner point) are vorton_index[cell_start[c] .. cell_start[c] +
 * cell_count[c]), in increasing index order.  Device pointers owned by the
 * tree, valid until it is rebuilt; *cells = leaf grid capacity. */
vg_status vg_tree_partition(const vg_tree *t, const uint32_t **cell_start,
                            const uint32_t **cell_count,
                            const uint32_t **vorton_index, size_t *cells);
/* VortonSim::DiffuseAndDissipateVorticityPSE (vortonSim.cpp:3616-3922) on
 * the vortons the tree was built from (they must not have moved since):
 * particle strength exchange over each cell and its 7 forward neighbours,
 * f = clamp(2 nu dt exp(-d^2/(4 size)^2), +-0.5), then dissipation
 * Omega -= nu dt Omega, in the order of the reference's threaded build (even
 * z slices, then odd; the result depends on that order and is reproduced
 * bit for bit).  angvel (strided) is read and updated in place; birth
 * (int32, strided, may be NULL) marks dead vortons (birth < 0), which
 * receive exchanges but neither initiate them nor dissipate.  Requires
 * nu dt <= 1 (the reference's assertion).
 * order VG_PSE_REFERENCE_ORDER reproduces the reference.  The reference's
 * in-place exchanges are a long serial chain per z slice, run here as
 * wavefronts of cells: correct but latency-bound.  This order reads back the
 * size of its transient fraction buffer, so the call synchronises stream s
 * once (it cannot be recorded into a graph).  VG_PSE_SIMULTANEOUS (no
 * reference counterpart) takes every exchange from the starting values, one
 * thread per vorton.  Its gap from the in-place order is second order in
 * the total exchange fraction a vorton sees, sum_j f_j, which grows with
 * the partner count.  At the demo's nu = 0.01, dt = 1/30 the rel-L2 gap is
 * 7.5e-6 for 8k colliding rings but 2.8e-3 for 1M vortons in 64 rings,
 * whose cells hold up to 69 vortons. */
enum {
    VG_PSE_REFERENCE_ORDER = 0,
    VG_PSE_SIMULTANEOUS = 1
};
vg_status vg_tree_pse_diffuse(const vg_tree *t, float *angvel, size_t ang_stride,
                              const int32_t *birth, size_t birth_stride,
                              float viscosity, float dt, int order, vg_stream s);

/* ------------------------------------------------------------ learned far-field correction */
/* No reference counterpart: the reference applies every cluster as a
 * monopole supervorton (vorton.h:161-179 on the clusters of
 * vortonSim.cpp:1337-1569).  A 16 -> 64 -> 64 -> 64 -> 3 GELU (tanh form)
 * network learns, per (query, cluster) pair, the part of the cluster's
 * velocity that its monopole and analytic first moment miss.  With h the
 * cube root of the layer's cell volume, S the cluster's weight sum,
 * x = (q - c)/h, r = |x|, e = x/r, O = Omega/S, Dt = D/(h S) (D the first
 * moment, row-major [b][e]), at its antisymmetric vector, u = Dt e,
 * v = Dt^T e, the inputs are the rotation invariants
 *   1/r, 1/r^2, |O|^2, (O.e)^2, |at|^2, (at.e)^2, |u|^2, |v|^2, (e.u)^2,
 *   |Dt|_F^2, (tr Dt)^2, O.at, (O.e)(e.u), O.u, O.v, at.(u x e)
 * and the outputs c weight B = r^-3 [O x e, at, u x e] (columns): the
 * correction (2/3)(size/2)^3 S/h^2 B c rotates with the field and reverses
 * with the vorticity.  Products run on BF16 MFMA with FP32 accumulation.
 * Parameters, compact order (VG_MLP_PARAMS floats): W1[16][64] b1[64]
 * W2[64][64] b2[64] W3[64][64] b3[64] W4[64][3] b4[3], W row-major
 * [in][out]. */
#define VG_MLP_INPUTS 16
#define VG_MLP_HIDDEN 64
#define VG_MLP_OUTPUTS 3
#define VG_MLP_PARAMS 9603
typedef struct vg_mlp vg_mlp;
/* Weights uniform in +-sqrt(6/(fan_in+fan_out)) from splitmix64(seed),
 * zero biases, Adam state cleared. */
vg_status vg_mlp_create(vg_mlp **m, uint64_t seed);
vg_status vg_mlp_destroy(vg_mlp *m);
/* Host arrays of VG_MLP_PARAMS floats; set_params also clears Adam state. */
vg_status vg_mlp_set_params(vg_mlp *m, const float *params);
vg_status vg_mlp_get_params(vg_mlp *m, float *params);
/* c[n][3] = net(x[n][16]) (device pointers). */
vg_status vg_mlp_forward(vg_mlp *m, const float *x, uint32_t n, float *y, vg_stream s);
/* Gradient (host, VG_MLP_PARAMS floats) of mean((B net(x) - t)^2) over the
 * n*3 components, and that loss (host); basis[n][3][3] (row i, column k;
 * NULL: identity) and t[n][3] on the device.  Blocks. */
vg_status vg_mlp_gradients(vg_mlp *m, const float *x, const float *basis, const float *t,
                           uint32_t n, float *grad, float *loss, vg_stream s);
/* steps Adam updates (lr, beta 0.9 / 0.999, eps 1e-8) on minibatches of
 * `batch` consecutive rows of (x, t), offsets (step * batch) mod
 * (n - batch + 1).  final_loss (host, may be NULL): loss of the last
 * minibatch before its update.  Blocks when final_loss is given. */
vg_status vg_mlp_train(vg_mlp *m, const float *x, const float *basis, const float *t,
                       uint32_t n, uint32_t batch, uint32_t steps, float lr, float *final_loss,
                       vg_stream s);
/* Training pairs from a tree: n random (grid point, accepted cluster)
 * pairs of the near-field treecode with radius near_k on the layers in
 * layer_mask (bit l = cluster layer l); x[n][16] the network inputs,
 * basis[n][3][3] the output vectors (may be NULL), target[n][3] = (direct
 * sum over the cluster's member vortons - its monopole - its first-moment
 * term) divided by (2/3)(size/2)^3 S/h^2, with (spread, circ).
 * meta (device, may be NULL): per pair the grid point offset and
 * layer * 2^28 + cluster point offset (0xffffffff for a pair left empty).
 * n_failed (host, may be NULL): pairs left zero because no cluster was
 * found in 256 draws.  Blocks. */
vg_status vg_tree_mlp_samples(const vg_tree *t, const vg_grid_desc *g, int near_k,
                              uint32_t layer_mask, float spread, float circ, uint32_t n,
                              uint64_t seed, float *x, float *basis, float *target,
                              uint32_t *meta, uint32_t *n_failed, vg_stream s);
/* Near-field treecode with first moments (vg_tree_velocity_grid_mac with
 * order 1) plus the network's correction for every applied cluster on the
 * layers in layer_mask.  layer_mask 0 leaves the plain treecode result. */
vg_status vg_tree_velocity_grid_mlp(const vg_tree *t, const vg_grid_desc *g, vg_mlp *m,
                                    int near_k, uint32_t layer_mask, float spread, float circ,
                                    int rsqrt_policy, float *out, vg_stream s);

/* ------------------------------------------------------------ batched effects */
/* Many independent small simulations ("effects") advanced together; no
 * reference counterpart (the reference runs one VortonSim per particle
 * group, vortonSim.cpp:4117-4256).  Each effect has its own vortons, tracers
 * and vorton radius, and a grid of cells^3 cells over its own bounding box
 * (vortons u tracers, padded by 3 radii as UpdateBoundingBox does).  A frame,
 * in the reference's order (particleSystemConfiguration.cpp:282-296):
 * tracers AssignVelocityFromField(previous grid) + Evolve; bounds; influence
 * tree; near-field treecode velocity at the grid points (near_k, order,
 * auto-mollified); vortons AssignVelocityFromField + Evolve.  Every stage is
 * one launch over all effects and nothing returns to the host, so a frame
 * can be recorded into a hipGraph (vg_batch_capture) and replayed with one
 * launch (vg_batch_launch). */
typedef struct vg_batch vg_batch;
typedef struct vg_batch_config {
    uint32_t cells;    /* cells per axis of every effect's grid (default 32): a power
                        * of two or 3 * 2^k; counts that leave a tree cell without
                        * a parent (e.g. 20, 36) are refused with VG_EINVAL.
                        * 0: every frame, each effect's grid is
                        * UniformGridGeometry::DefineShape(vorton count, padded box,
                        * power-of-two cells) as UpdateBoundingBox calls it
                        * (uniformGrid.h:259-322, vortonSim.cpp:1089), computed on
                        * the device; the walk then takes (effect, brick) items
                        * from a device work queue, largest effects first */
    int near_k;        /* near-field treecode radius (default 2); -1: the reference
                        * opening rule instead (ComputeVelocityAtGridpoints of the
                        * TREE technique, vortonSim.cpp:1787-1905, 2223-2293), as
                        * vg_sim with tree_near_k < 0; order and expand_from are
                        * then unused */
    int order;         /* cluster order: 0 monopole, 1 dipole (default 1) */
    int rsqrt_policy;  /* VG_RSQRT_* (default ACCURATE) */
    int expand_from;   /* local expansions from this layer (default -1: off) */
    uint32_t max_cells; /* cells == 0: at most this many cells per axis, a power of two
                         * in [4, 256] (default 64); storage is sized for it.  A
                         * shape beyond it on an axis is limited to it there and
                         * counted (vg_batch_clamped) */
} vg_batch_config;
vg_status vg_batch_config_default(vg_batch_config *cfg);
/* Host arrays of `effects` entries; tracer_counts may be NULL (none). */
vg_status vg_batch_create(vg_batch **b, uint32_t effects, const uint32_t *vorton_counts,
                          const uint32_t *tracer_counts, const float *vorton_radii,
                          const vg_batch_config *cfg);
vg_status vg_batch_destroy(vg_batch *b);
/* Device buffers owned by the batch, packed xyz, effects back to back in
 * order: vorton positions, angular velocities (vorticity / 2), velocities,
 * tracer positions.  Any pointer may be NULL. */
vg_status vg_batch_buffers(vg_batch *b, float **vorton_pos, float **vorton_angvel,
                           float **vorton_vel, float **tracer_pos, size_t *n_vortons,
                           size_t *n_tracers);
/* One frame, launched eagerly. */
vg_status vg_batch_step(vg_batch *b, float dt, vg_stream s);
/* Records one frame with time step dt into a hipGraph (s must be a created
 * stream); vg_batch_launch replays it.  Changing the launch shapes drops the
 * recording. */
vg_status vg_batch_capture(vg_batch *b, float dt, vg_stream s);
vg_status vg_batch_launch(vg_batch *b, vg_stream s);
/* Grid descriptor (host copy) and device field of one effect's last frame,
 * and its mollification factors.  Synchronise the device. */
vg_status vg_batch_grid(const vg_batch *b, uint32_t effect, vg_grid_desc *g, const float **field);
vg_status vg_batch_mollify(const vg_batch *b, uint32_t effect, float *spread, float *circ);
/* The velocity grids of the current particles without moving them: bounds,
 * geometry, trees and the walk (the middle of vg_batch_step; equal to a
 * step with dt = 0).  For evaluating methods, e.g. by the autotuner. */
vg_status vg_batch_solve(vg_batch *b, vg_stream s);
/* Number of (effect, frame) grid definitions limited by max_cells since
 * creation.  Blocks. */
vg_status vg_batch_clamped(const vg_batch *b, uint32_t *count);
/* Reorders each effect's tracers into the cell order of its current velocity
 * grid (the grid the next frame's tracer pass interpolates), by a stable
 * counting sort; every tracer stays in its effect's range.  The tracer
 * order carries no meaning in the reference (Particles::KillParticle moves
 * the last particle into the hole, particle.cpp:103), and cell order makes
 * the tracer pass's grid gathers fall on shared cache lines.  Needs a
 * computed frame (VG_ESTATE otherwise).  The first call allocates the sort
 * buffers (~28 B per tracer + 8 B per grid point): make it outside graph
 * capture. */
vg_status vg_batch_sort_tracers(vg_batch *b, vg_stream s);
/* Changes the walk's method (near-field radius, cluster order, first
 * expanded layer) for later frames; drops a recorded frame. */
vg_status vg_batch_set_method(vg_batch *b, int near_k, int order, int expand_from);
/* Launch shapes: treecode waves per workgroup {1, 2, 4}; vorton-advection
 * workgroup size (multiple of 64, <= 1024) and vortons per thread {1, 2, 4}
 * (the tracers move in one fused advection + bounds pass). */
vg_status vg_batch_set_launch(vg_batch *b, int walk_waves, int advect_block, int advect_items);
vg_status vg_batch_get_launch(const vg_batch *b, int *walk_waves, int *advect_block,
                              int *advect_items);
/* Times every candidate launch shape on the live batch (after at least one
 * vg_batch_step) and keeps the fastest; report (may be NULL) receives one
 * line per trial.  Blocks.  Results of later frames do not depend on the
 * shapes chosen.  The method (near_k, expand_from) is chosen against an
 * accuracy target by vortegrid_amd/autotune.py, which drives vg_batch_solve
 * and this call. */
vg_status vg_batch_autotune(vg_batch *b, vg_stream s, char *report, size_t cap);

/* ------------------------------------------------------------ simulation */
/* vg_sim = VortonSim (vortonSim.h:137-623) with device-resident state:
 * grid template, influence tree, multigrid pyramids and velocity grid. */
typedef struct vg_sim vg_sim;
typedef struct vg_sim_config {
    int technique;       /* VG_TECH_*; reference default is POISSON */
    int rsqrt_policy;    /* VG_RSQRT_* */
    int auto_mollify;    /* 1: ENABLE_AUTO_MOLLIFICATION for TREE */
    uint32_t mg_steps;   /* SOR steps per level (reference: 16) */
    int boundary_copy_odd; /* 1: odd-x face points copy x-1 (reference) */
    int poisson_solver;  /* 0: reference cascade, 1: V-cycles to mg_tol */
    double mg_tol;
    uint32_t mg_max_cycles;
    /* TREE technique grid evaluation: tree_near_k < 0 keeps the reference
     * opening rule; tree_near_k >= 0 selects the near-field treecode with
     * that near radius and cluster order tree_order (0 monopole, 1 dipole). */
    int tree_near_k;
    int tree_order;
    /* near-field treecode: first layer applied through brick-centred local
     * expansions (see vg_tree_velocity_grid_mac); < 1 disables. */
    int tree_expand_from;
} vg_sim_config;
vg_status vg_sim_config_default(vg_sim_config *cfg);
vg_status vg_sim_create(vg_sim **sim, const vg_sim_config *cfg, vg_stream s);
vg_status vg_sim_destroy(vg_sim *sim);
/* Learned far-field correction for the TREE technique (no reference
 * counterpart): with tree_near_k >= 0 and m != NULL the velocity grid is
 * vg_tree_velocity_grid_mlp(near_k, layer_mask) instead of the near-field
 * treecode of order tree_order; m == NULL restores the plain treecode.  The
 * network is not owned and must outlive its use. */
vg_status vg_sim_set_mlp(vg_sim *sim, vg_mlp *m, uint32_t layer_mask);
/* VortonSim::SetVortons (vortonSim.h:329): keeps the device view. */
vg_status vg_sim_set_vortons(vg_sim *sim, const vg_vortons *v);
/* VortonSim::FindBoundingBox (vortonSim.cpp:1001-1008). */
vg_status vg_sim_find_bounding_box(vg_sim *sim);
/* VortonSim::FindBoundingBox together with PclOpFindBoundingBox::Operate
 * on one more particle group (the tracers, pclOpFindBoundingBox.cpp:
 * 141-148), in one pass and one host wait: the sim's box becomes its
 * vortons' box, and mn/mx (host, in/out) grow by the n strided positions
 * at pos.  Equal to vg_find_bounding_box + vg_sim_find_bounding_box. */
vg_status vg_sim_find_bounding_boxes(vg_sim *sim, const float *pos, size_t stride, size_t n,
                                     float mn[3], float mx[3]);
/* VortonSim::UpdateBoundingBox (vortonSim.cpp:1027-1096). */
vg_status vg_sim_update_bounding_box(vg_sim *sim, const float mn[3],
                                     const float mx[3], int final_update);
/* VortonSim::Update -> UpdateVortexParticleMethod (vortonSim.cpp:4117-4256),
 * velocity-from-vorticity part: influence tree, vorticity splat (POISSON),
 * velocity grid, then AssignVelocityFromField(vortons, grid, 1) into
 * vel_out (strided like the vortons; NULL skips it). */
vg_status vg_sim_update(vg_sim *sim, float dt, uint32_t frame, float *vel_out,
                        size_t vel_stride);
/* VortonSim::GetVelocityGrid (vortonSim.h:334): device pointer valid until
 * the next update. */
vg_status vg_sim_velocity_grid(const vg_sim *sim, vg_grid_desc *g,
                               const float **field);
/* VortonSim::StretchAndTiltVortons (vortonSim.cpp:2748-2820) on the last
 * update's velocity grid and the vortons' positions: angvel (strided by
 * ang_stride, normally the vortons' own angular velocities) is updated in
 * place, processed in the cell order of the update's influence tree (the
 * vorton count must not have changed since).  The reference runs it inside
 * Update for the vortex particle method (:4208-4215). */
vg_status vg_sim_stretch_tilt(vg_sim *sim, float dt, float *angvel,
                              size_t ang_stride);
/* VortonSim::DiffuseAndDissipateVorticityPSE (vortonSim.cpp:3616-3922) on
 * the update's partition: vg_tree_pse_diffuse on the sim's tree.  Call
 * after vg_sim_update (and vg_sim_stretch_tilt), before the vortons move;
 * the reference runs it inside Update when its investigation term is ALL or
 * VISCOUS_DIFFUSION (:4238-4241). */
vg_status vg_sim_pse_diffuse(vg_sim *sim, float viscosity, float dt, float *angvel,
                             size_t ang_stride, const int32_t *birth,
                             size_t birth_stride, int order);
/* GetVectorPotentialMultiGrid()[layer] / GetNegativeVorticityMultiGrid()[layer]
 * (vortonSim.h:338-343) after a POISSON update: layer 0 is the full-size
 * field, layers >= 1 the cascade's coarse layers (the V-cycle solver keeps
 * only layer 0).  depth (may be NULL) receives the layer count (0 before a
 * POISSON update).  Device pointers, valid until the next update. */
vg_status vg_sim_multigrid(const vg_sim *sim, int field, uint32_t layer, vg_grid_desc *g,
                           const float **data, uint32_t *depth);
/* GetVectorPotentialMultiGrid()[0] / GetNegativeVorticityMultiGrid()[0]. */
vg_status vg_sim_vector_potential(const vg_sim *sim, vg_grid_desc *g,
                                  const float **field);
vg_status vg_sim_negative_vorticity(const vg_sim *sim, vg_grid_desc *g,
                                    const float **field);
vg_status vg_sim_grid_template(const vg_sim *sim, vg_grid_desc *g);
vg_status vg_sim_bounds(const vg_sim *sim, float mn[3], float mx[3]);
vg_status vg_sim_tree(const vg_sim *sim, const vg_tree **t);
/* Mollification factors used by the last update. */
vg_status vg_sim_spreading(const vg_sim *sim, float *spread, float *circ);
/* FP64 true relative residual of the last Poisson solve (POISSON only). */
vg_status vg_sim_poisson_residual(const vg_sim *sim, double *rel);

#ifdef __cplusplus
}
#endif
#endif
