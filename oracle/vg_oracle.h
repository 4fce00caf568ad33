/* vg_oracle.h -- CPU restatement of VorteGrid's velocity-from-vorticity path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation in vortegrid_amd/: only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it.  The product never links it.
 *
 * Every function restates, as plain C (FP32, no FMA contraction, serial loop
 * order), the reference routine cited beside it.  Citations are relative to
 * Samples/MjgIntelFluidDemo_Part20/ of mijagourlay/VorteGrid.
 *
 * Parity status of each function (what pins it) is listed in oracle/README.md.
 */
#ifndef VG_ORACLE_H
#define VG_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Grid geometry: UniformGridGeometry (Core/SpatialPartition/uniformGrid.h:144-860). */
typedef struct vgo_grid {
    float    min[3];      /* mMinCorner */
    float    extent[3];   /* mGridExtent */
    float    cell[3];     /* mCellExtent (spacing) */
    float    cpe[3];      /* mCellsPerExtent */
    uint32_t np[3];       /* mNumPoints */
} vgo_grid;

/* ---- math primitives ---- */
float    vgo_finvsqrtf(float v);                         /* Core/Math/math.h:206-216 */
uint32_t vgo_nearest_pow2(uint32_t v);                   /* Core/Math/math.h:149-175 */

/* ---- pair kernels (VortonFluid/vorton.h:118-179) ----
 * rsqrt_policy 0 = reference fast inverse square root, 1 = IEEE 1/sqrtf. */
void vgo_accumulate_velocity(float acc[3], const float q[3], const float p[3],
                             const float om[3], float size, float spread,
                             float circ, int rsqrt_policy);
void vgo_accumulate_vecpot(float acc[3], const float q[3], const float p[3],
                           const float om[3], float size, float spread,
                           float circ, int rsqrt_policy);

/* ---- direct sums (VortonFluid/vortonSim.cpp:1113-1158) ----
 * pos/om/size are packed arrays (3, 3, 1 floats per vorton), om = angular
 * velocity (half vorticity).  Queries q are packed xyz.  out is packed xyz. */
void vgo_direct_velocity(const float *pos, const float *om, const float *size,
                         size_t n, const float *q, size_t nq, float spread,
                         float circ, int rsqrt_policy, float *out, int nthreads);
void vgo_direct_vecpot(const float *pos, const float *om, const float *size,
                       size_t n, const float *q, size_t nq, float spread,
                       float circ, int rsqrt_policy, float *out, int nthreads);
/* FP64 accuracy oracle: same mollified law, exact 1/sqrt, double accumulate. */
void vgo_direct_velocity_f64(const float *pos, const float *om, const float *size,
                             size_t n, const float *q, size_t nq, float spread,
                             float circ, double *out, int nthreads);

/* ---- geometry (uniformGrid.h:259-322, 361-376, 794-809) ---- */
void vgo_define_shape(vgo_grid *g, size_t n_elements, const float vmin[3],
                      const float vmax[3], int pow2);
void vgo_decimate(vgo_grid *dst, const vgo_grid *src, uint32_t decimation);
size_t vgo_grid_capacity(const vgo_grid *g);
/* Grid point position with the (1-2eps) spacing nudge used by the grid-point
 * query loops (vortonSim.cpp:2064-2065, 2229-2230). */
void vgo_grid_query_point(const vgo_grid *g, uint32_t ix, uint32_t iy,
                          uint32_t iz, float out[3]);
/* IndicesOfPosition (uniformGrid.h:536-574): truncation toward zero. */
void vgo_indices_of_position(const vgo_grid *g, const float p[3], uint32_t idx[3]);

/* ---- bounding box (vortonSim.cpp:1001-1096, pclOpFindBoundingBox.cpp) ---- */
void vgo_find_bbox(const float *pos, size_t stride, size_t n, float mn[3],
                   float mx[3]);
/* UpdateBoundingBox(min,max,bFinal=true) + DefineShape(N, ..., pow2=true). */
void vgo_finalize_bbox(float mn[3], float mx[3], float vorton_radius,
                       size_t n_vortons, vgo_grid *grid_template);

/* ---- nested grid (Core/SpatialPartition/nestedGrid.h:69-391) ---- */
#define VGO_MAX_LAYERS 24
typedef struct vgo_nested {
    uint32_t depth;
    vgo_grid layer[VGO_MAX_LAYERS];
    uint32_t dec[VGO_MAX_LAYERS][3];   /* decimations of parent layer l */
} vgo_nested;
void vgo_nested_init(vgo_nested *ng, const vgo_grid *src);

/* ---- influence tree (vortonSim.cpp:1192-1624) ----
 * Per layer, per grid point: cluster position (3), angular velocity (3),
 * size (1).  Caller allocates with vgo_tree_alloc. */
typedef struct vgo_tree {
    vgo_nested ng;
    float *pos[VGO_MAX_LAYERS];
    float *om[VGO_MAX_LAYERS];
    float *size[VGO_MAX_LAYERS];
    /* First moment of each cluster about its centroid (not in the
     * reference): dip[9*c + 3*b + e] = sum_i Omega_i[b] * (y_i - c)[e]. */
    float *dip[VGO_MAX_LAYERS];
    /* Leaf cell lists: vortons of leaf cell c (point offset) are
     * sorted_idx[cell_start[c] .. cell_start[c] + cell_count[c]) in index
     * order. */
    uint32_t *cell_start, *cell_count, *sorted_idx;
    size_t n;
} vgo_tree;
int  vgo_tree_build(vgo_tree *t, const vgo_grid *grid_template,
                    const float *pos, const float *om, const float *size,
                    size_t n);
void vgo_tree_free(vgo_tree *t);

/* ComputeVelocity_Tree / ComputeVectorPotential_Tree from the root
 * (vortonSim.cpp:1650-1905), recursive, serial accumulation order. */
void vgo_tree_velocity(const vgo_tree *t, const float q[3], float spread,
                       float circ, int rsqrt_policy, float out[3]);
void vgo_tree_vecpot(const vgo_tree *t, const float q[3], float spread,
                     float circ, int rsqrt_policy, float out[3]);

/* Near-field treecode at every grid point (this build's accurate mode; no
 * reference counterpart).  Grid points are taken in 4x4x4 bricks; for each
 * brick and layer the near box is the brick's cells +-near_k, clipped to
 * the children of the parent layer's near box.  Children outside the near
 * box are applied as clusters (monopole, plus the first-moment correction
 * when order == 1); leaf cells inside the leaf near box are summed over
 * their vortons.  Per-lane sums follow the GPU order: layers from the top,
 * cells z,y,x, one Kahan-compensated partial per layer. */
int vgo_mac_velocity_grid(const vgo_tree *t, const vgo_grid *g, const float *pos,
                          const float *om, const float *size, int near_k,
                          int order, float spread, float circ, int rsqrt_policy,
                          float *out, int nthreads);
/* Same, restricted to z-brick rows [bz_begin, bz_end) (timing samples). */
int vgo_mac_velocity_grid_bricks(const vgo_tree *t, const vgo_grid *g, const float *pos,
                                 const float *om, const float *size, int near_k, int order,
                                 float spread, float circ, int rsqrt_policy, float *out,
                                 int nthreads, uint32_t bz_begin, uint32_t bz_end);

/* Same with clusters of layers >= expand_from applied through a local
 * expansion about each brick's centre (no reference counterpart; monopole
 * to second order, first moment to first order, float64 tensors);
 * expand_from < 1 or > depth-2 disables it.
 * All three return 0, or -1 without touching out when the tree's grid leaves
 * a cell without a parent cluster (vgo_nested_has_orphans): the walk from the
 * upper layers would never reach that cell's vortons. */
int vgo_mac_velocity_grid_ex(const vgo_tree *t, const vgo_grid *g, const float *pos,
                             const float *om, const float *size, int near_k, int order,
                             int expand_from, float spread, float circ, int rsqrt_policy,
                             float *out, int nthreads, uint32_t bz_begin, uint32_t bz_end);
/* Same on the bricks with x index in [bx_begin, bx_end) and z index in
 * [bz_begin, bz_end) only (timing samples). */
int vgo_mac_velocity_grid_box(const vgo_tree *t, const vgo_grid *g, const float *pos,
                              const float *om, const float *size, int near_k, int order,
                              int expand_from, float spread, float circ, int rsqrt_policy,
                              float *out, int nthreads, uint32_t bx_begin, uint32_t bx_end,
                              uint32_t bz_begin, uint32_t bz_end);
/* 1 when some layer l has more cells on an axis than dec[l] x its parent's
 * (an odd count above 3 halved, e.g. 5 cells under 2 parents): the last
 * cell then has no parent in AggregateClusters (vortonSim.cpp:1337-1569).
 * Power-of-two grids, the only kind the reference builds, never do. */
int vgo_nested_has_orphans(const vgo_nested *ng);

/* The local expansion of one cluster about cb: out = c0[3], c1[3][3],
 * c2[3][3][3] (see vgo_mac_velocity_grid_ex). */
void vgo_local_expansion(const float cb[3], const float cp[3], const float w[3], float size,
                         const float D[9], int order, float spread, double out[39]);

/* Auto-mollification (vortonSim.cpp:2458-2477). */
void vgo_auto_mollify(const vgo_grid *vel_grid, float vorton_radius,
                      float *spread, float *circ);

/* Velocity at every grid point (vortonSim.cpp:2223-2293): technique 0 =
 * direct, 1 = tree (falls back to direct when depth==1). */
void vgo_velocity_grid(const vgo_tree *t, const vgo_grid *g, const float *pos,
                       const float *om, const float *size, size_t n,
                       int technique, float spread, float circ,
                       int rsqrt_policy, float *out, int nthreads);

/* Boundary vector potential (vortonSim.cpp:2055-2138, boundariesOnly=true):
 * even x computed by tree (or direct when depth==1), odd x copies x-1.
 * copy_odd=0 computes every face point (accuracy mode).  Interior of out is
 * left untouched. */
void vgo_vecpot_boundary(const vgo_tree *t, const vgo_grid *g,
                         const float *pos, const float *om, const float *size,
                         size_t n, float spread, float circ, int rsqrt_policy,
                         int copy_odd, float *out, int nthreads);

/* ---- grid transfer (uniformGrid.h:1117-1342, 1453-1718) ---- */
void vgo_interpolate(const vgo_grid *g, const float *field, const float p[3],
                     float out[3]);
void vgo_accumulate(const vgo_grid *g, float *field, int ncomp,
                    const float p[3], const float *item);
/* PopulateVorticityGridFromVortons (vortonSim.cpp:2372-2436). */
void vgo_splat_vorticity(const vgo_grid *g, const float *pos, const float *om,
                         size_t n, float vorton_radius, float scale,
                         float *out);
/* Particles::PopulateVelocityGrid (pclOpPopulateVelocityGrid.cpp:33-87). */
void vgo_populate_velocity_grid(const vgo_grid *g, const float *pos,
                                const float *vel, size_t n, float *out);
/* DownSample FASTER_LESS_ACCURATE (injection) and SLOWER_MORE_ACCURATE. */
void vgo_downsample(const vgo_grid *lo, float *lo_field, const vgo_grid *hi,
                    const float *hi_field, int accurate);
/* UpSample (serial semantics); interior_only selects INTERIOR_ONLY. */
void vgo_upsample(const vgo_grid *hi, float *hi_field, const vgo_grid *lo,
                  const float *lo_field, int interior_only);

/* ---- Poisson (uniformGridMath.cpp:1063-1405) ---- */
/* One red (0) or black (1) zebra pass, Dirichlet, SOR relax 1.72. */
void vgo_sor_pass(const vgo_grid *g, float *soln, const float *lap, int color);
void vgo_solve_vector_poisson(const vgo_grid *g, float *soln, const float *lap,
                              size_t num_steps);
/* Multigrid cascade of VortonSim::ComputeVectorPotential (vortonSim.cpp:
 * 2523-2608).  a0 holds boundary values on entry (interior ignored, set to 0);
 * negw0 holds the negative vorticity.  Returns solution in a0. */
void vgo_mg_cascade(const vgo_grid *g, float *a0, const float *negw0,
                    size_t steps_per_level);
/* ComputeJacobian + ComputeCurlFromJacobian (uniformGridMath.cpp:207-451). */
void vgo_curl(const vgo_grid *g, const float *a, float *curl);
/* ComputeJacobian (uniformGridMath.cpp:304-451): jac[9 o + 3 a + b] =
 * d v_b / d a (Mat33 members x, y, z hold d/dx, d/dy, d/dz). */
void vgo_jacobian(const vgo_grid *g, const float *v, float *jac);
/* StretchAndTiltVortons (vortonSim.cpp:2748-2820): per vorton, trilinear
 * Jacobian at the position (UniformGrid<Mat33>::Interpolate,
 * uniformGrid.h:1117-1148), s = Omega * J (mat33.h:131-145: s_a = J.a . Omega),
 * Omega += 0.5 s dt.  No-op when any grid extent is zero. */
void vgo_stretch_tilt(const vgo_grid *g, const float *jac, const float *pos,
                      size_t pos_stride, float *angvel, size_t ang_stride,
                      size_t n, float dt);
/* exp(x) by Cody-Waite reduction x = k ln2 + r (ln2 split into a 16-bit
 * high part and a low part), the degree-7 Taylor polynomial of exp(r) in
 * Horner form, and 2^k built from the exponent bits; 0 below -87.  The
 * PSE kernel's distance law, evaluated identically on the CPU and GPU. */
float vgo_pse_exp(float x);
/* VortonSim::DiffuseAndDissipateVorticityPSE (vortonSim.cpp:3616-3922) in
 * the order of the reference's threaded build (USE_TBB, :3855-3860): all
 * even z cell slices, then all odd ones; within a slice cells in (y, x)
 * order, within a cell vortons in index order (the partition of
 * PartitionVortons, :3578-3607, particle.cpp:398-435).  Each live vorton h
 * exchanges with the later vortons of its cell, then with those of the 7
 * forward cells (+x, +y, +x+y, +z, +x+z, +y+z, +x+y+z, :3711-3728):
 * f = clamp(2 nu dt exp(-d^2 / (4 size_h)^2), +-0.5), e = f (Omega_h -
 * Omega_t), Omega_h -= e, Omega_t += e; then Omega_h -= nu dt Omega_h.
 * pos, angvel: packed n x 3 (angvel in/out); size: n diameters or NULL
 * for uniform_size; alive: n flags or NULL (all alive).  Merging is off in
 * the reference build (ENABLE_MERGING_VORTONS undefined). */
void vgo_pse_diffuse(const vgo_grid *g, const float *pos, float *angvel,
                     const float *size, float uniform_size,
                     const unsigned char *alive, size_t n, float viscosity,
                     float dt);
/* The same pairs and formulas applied simultaneously (no reference
 * counterpart): every exchange e = f (Omega_h - Omega_t) uses the starting
 * values, vorton i sums, in this order, the exchanges it receives from live
 * vortons of the 7 backward cells (offsets -fwd[0..6], each cell in index
 * order), from live earlier vortons of its own cell, then (if alive) minus
 * those it makes with the later vortons of its cell and the 7 forward
 * cells; Omega_i += sum, then Omega_i -= nu dt Omega_i if alive.  Differs
 * from the in-place order at second order in the exchange fractions a
 * vorton accumulates over its partners. */
void vgo_pse_diffuse_simultaneous(const vgo_grid *g, const float *pos, float *angvel,
                                  const float *size, float uniform_size,
                                  const unsigned char *alive, size_t n,
                                  float viscosity, float dt);
/* UniformGrid<Vec3>::GenerateBrickOfBytes (uniformGridDiagnostics.cpp:
 * 209-314), data part: planes[k * cap + o] for k = x, y, z, magnitude;
 * stats = {min x, y, z, max x, y, z, magnitude max} with ComputeStatistics'
 * all-components Min2/Max2 scan (uniformGrid.h:1095-1107, macros.h:61-64). */
void vgo_brick_of_bytes(const vgo_grid *g, const float *field, unsigned char *planes,
                        float stats[7]);
/* FP64 residual ||f - L u|| / ||f|| over interior points, 7-point Laplacian. */
double vgo_residual_rel(const vgo_grid *g, const float *u, const float *f);
/* The same for the compensated pair u = hi + lo (lo may be NULL). */
double vgo_residual_rel_pair(const vgo_grid *g, const float *hi, const float *lo,
                             const float *f);

/* ---- particle operators ---- */
/* AssignParticleVelocityFromField_Slice (pclOpAdvect.cpp:42-81). */
void vgo_assign_velocity_from_field(const vgo_grid *g, const float *field,
                                    float *pos, float *vel, size_t stride,
                                    size_t n, float gain);
/* EvolveParticlesSlice (pclOpEvolve.cpp:36-55). */
void vgo_evolve(float *pos, const float *vel, size_t stride, size_t n,
                float dt);

#ifdef __cplusplus
}
#endif
#endif
