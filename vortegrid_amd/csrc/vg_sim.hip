// vg_sim.hip -- VortonSim's velocity-from-vorticity frame with device-resident
// state (vortonSim.h:137-623).  Orchestration only: every numeric step is one
// of the operators in vg_direct / vg_tree / vg_grid / vg_multigrid.
//
// Frame (UpdateVortexParticleMethod, vortonSim.cpp:4117-4256, velocity part):
//   CreateInfluenceTree                          -> vg_tree_build
//   POISSON: PopulateVorticityGridFromVortons(-1) -> vg_tree_splat_vorticity
//            ComputeVectorPotential               -> faces by treecode + MG
//            Jacobian + curl                      -> vg_curl
//   TREE:    auto-mollify + treecode at grid points
//   DIRECT:  direct velocity at vortons + PopulateVelocityGrid
//   AssignVelocityFromField(vortons, grid, 1)    (not for DIRECT, whose
//                                                 vortons keep the direct
//                                                 velocity)
//   StretchAndTiltVortons                        -> vg_sim_stretch_tilt
//   DiffuseAndDissipateVorticityPSE              -> vg_sim_pse_diffuse
//                                                 (both called after the update)
#include <cstring>

#include "vg_internal.h"

namespace vg {
vg_status curl(const vg_grid_desc &g, const float *a, float *out, hipStream_t s);
vg_status residual_norm(const vg_grid_desc &g, const float *u, const float *f, double *rel,
                        hipStream_t s);
vg_status stretch_tilt_fused(const vg_grid_desc &g, const float *vel, const float *pos,
                             size_t pos_stride, float *angvel, size_t ang_stride, size_t n,
                             float dt, const uint32_t *order, hipStream_t s);
vg_status tree_checked_view(const vg_tree *t, TreeView *v);
uint32_t mg_cascade_depth(const vg_multigrid *m);
vg_status find_bounding_boxes(uint32_t count, const float *const *pos, const size_t *stride,
                              const size_t *n, float *mn, float *mx, hipStream_t s);
vg_status pse_diffuse(const vg_tree *t, float *angvel, size_t ang_stride, const int32_t *birth,
                      size_t birth_stride, float viscosity, float dt, int order, hipStream_t s);
}  // namespace vg

struct vg_sim {
    vg_sim_config cfg;
    hipStream_t stream = nullptr;
    vg_vortons v{};
    bool have_vortons = false;
    float mn[3], mx[3];
    vg_grid_desc tmpl{};
    vg_tree *tree = nullptr;
    vg_multigrid *mg = nullptr;
    float *vel_grid = nullptr, *a0 = nullptr, *negw = nullptr, *vort_vel = nullptr;
    size_t vel_cap = 0, a0_cap = 0, negw_cap = 0, vort_cap = 0;
    vg_grid_desc vel_desc{};
    float spread = 1.f, circ = 1.f;
    double residual = -1.0;
    vg_mlp *mlp = nullptr;     // learned correction (not owned)
    uint32_t mlp_layers = 0;
};

using namespace vg;

static vg_status grow_f(float **p, size_t *cap, size_t need) {
    if (need <= *cap && *p) return VG_OK;
    if (*p) VG_HIP(hipFree(*p));
    *p = nullptr;
    *cap = 0;
    VG_HIP(hipMalloc((void **)p, need * sizeof(float)));
    *cap = need;
    return VG_OK;
}

VG_API vg_status vg_sim_config_default(vg_sim_config *c) {
    VG_REQUIRE(c, "config is null");
    memset(c, 0, sizeof *c);
    c->technique = VG_TECH_POISSON;
    c->rsqrt_policy = VG_RSQRT_REFERENCE;
    c->auto_mollify = 1;
    c->mg_steps = 16;
    c->boundary_copy_odd = 1;
    c->poisson_solver = 0;
    c->mg_tol = 1e-6;
    c->mg_max_cycles = 30;
    c->tree_near_k = -1;
    c->tree_order = 1;
    c->tree_expand_from = -1;
    return VG_OK;
}

VG_API vg_status vg_sim_create(vg_sim **out, const vg_sim_config *cfg, vg_stream s) {
    VG_REQUIRE(out, "out is null");
    vg_sim *sim = new vg_sim();
    if (cfg) sim->cfg = *cfg;
    else vg_sim_config_default(&sim->cfg);
    if (sim->cfg.technique < VG_TECH_DIRECT || sim->cfg.technique > VG_TECH_POISSON) {
        delete sim;
        set_error("bad technique");
        return VG_EINVAL;
    }
    sim->stream = (hipStream_t)s;
    for (int a = 0; a < 3; ++a) { sim->mn[a] = FLT_MAX; sim->mx[a] = -FLT_MAX; }
    vg_status st = vg_tree_create(&sim->tree);
    if (st == VG_OK) st = vg_mg_create(&sim->mg);
    if (st != VG_OK) {
        vg_tree_destroy(sim->tree);
        delete sim;
        return st;
    }
    *out = sim;
    return VG_OK;
}

VG_API vg_status vg_sim_destroy(vg_sim *sim) {
    if (!sim) return VG_OK;
    vg_tree_destroy(sim->tree);
    vg_mg_destroy(sim->mg);
    (void)hipFree(sim->vel_grid);
    (void)hipFree(sim->a0);
    (void)hipFree(sim->negw);
    (void)hipFree(sim->vort_vel);
    delete sim;
    return VG_OK;
}

VG_API vg_status vg_sim_set_mlp(vg_sim *sim, vg_mlp *m, uint32_t layer_mask) {
    VG_REQUIRE(sim, "null argument");
    sim->mlp = m;
    sim->mlp_layers = layer_mask;
    return VG_OK;
}

VG_API vg_status vg_sim_set_vortons(vg_sim *sim, const vg_vortons *v) {
    VG_REQUIRE(sim && v, "null argument");
    VG_REQUIRE(v->n == 0 || (v->pos && v->angvel && v->stride >= 3), "bad vorton view");
    VG_REQUIRE(v->n == 0 || v->size || v->uniform_size > 0.f, "vortons need a size");
    sim->v = *v;
    sim->have_vortons = true;
    return VG_OK;
}

static vg_status vorton_radius(vg_sim *sim, float *r) {
    if (sim->v.size) {
        float sz = 0.f;
        VG_HIP(hipMemcpyAsync(&sz, sim->v.size, sizeof(float), hipMemcpyDeviceToHost, sim->stream));
        VG_HIP(hipStreamSynchronize(sim->stream));
        *r = sz * 0.5f;
    } else {
        *r = sim->v.uniform_size * 0.5f;
    }
    return VG_OK;
}

VG_API vg_status vg_sim_find_bounding_box(vg_sim *sim) {
    VG_REQUIRE(sim && sim->have_vortons, "no vortons set");
    for (int a = 0; a < 3; ++a) { sim->mn[a] = FLT_MAX; sim->mx[a] = -FLT_MAX; }
    return vg_find_bounding_box(sim->v.pos, sim->v.stride, sim->v.n, sim->mn, sim->mx,
                                sim->stream);
}

VG_API vg_status vg_sim_find_bounding_boxes(vg_sim *sim, const float *pos, size_t stride,
                                            size_t n, float mn[3], float mx[3]) {
    VG_REQUIRE(sim && sim->have_vortons, "no vortons set");
    VG_REQUIRE(mn && mx && (n == 0 || (pos && stride >= 3)), "bad particle array");
    float bmn[6], bmx[6];
    for (int a = 0; a < 3; ++a) {
        bmn[a] = FLT_MAX;
        bmx[a] = -FLT_MAX;
        bmn[3 + a] = mn[a];
        bmx[3 + a] = mx[a];
    }
    const float *arrays[2] = {sim->v.pos, pos};
    const size_t strides[2] = {sim->v.stride, stride}, counts[2] = {sim->v.n, n};
    VG_TRY(find_bounding_boxes(2, arrays, strides, counts, bmn, bmx, sim->stream));
    for (int a = 0; a < 3; ++a) {
        sim->mn[a] = bmn[a];
        sim->mx[a] = bmx[a];
        mn[a] = bmn[3 + a];
        mx[a] = bmx[3 + a];
    }
    return VG_OK;
}

VG_API vg_status vg_sim_update_bounding_box(vg_sim *sim, const float mn[3], const float mx[3],
                                            int final_update) {
    VG_REQUIRE(sim && mn && mx, "null argument");
    for (int a = 0; a < 3; ++a) {
        sim->mn[a] = mn[a] < sim->mn[a] ? mn[a] : sim->mn[a];
        sim->mx[a] = mx[a] > sim->mx[a] ? mx[a] : sim->mx[a];
    }
    if (!final_update) return VG_OK;
    if (!sim->have_vortons || sim->v.n == 0) {
        const float z[3] = {0.f, 0.f, 0.f};
        define_shape(&sim->tmpl, 0, z, z, 1);
        return VG_OK;
    }
    float r = 0.f;
    VG_TRY(vorton_radius(sim, &r));
    return vg_finalize_bounding_box(sim->mn, sim->mx, r, sim->v.n, &sim->tmpl);
}

VG_API vg_status vg_sim_update(vg_sim *sim, float dt, uint32_t frame, float *vel_out,
                               size_t vel_stride) {
    (void)dt;
    (void)frame;
    VG_REQUIRE(sim && sim->have_vortons, "no vortons set");
    VG_REQUIRE(!vel_out || vel_stride >= 3, "bad velocity stride");
    hipStream_t s = sim->stream;
    const vg_grid_desc &g = sim->tmpl;
    sim->vel_desc = g;  // mVelGrid.CopyShape( mGridTemplate )
    const size_t cap = capacity(g);
    if (sim->v.n == 0 || cap == 0) {  // no vortons, no velocity (vortonSim.cpp:2672-2675)
        if (cap) {
            VG_TRY(grow_f(&sim->vel_grid, &sim->vel_cap, 3 * cap));
            VG_TRY(zero_fill(sim->vel_grid, 3 * cap * sizeof(float), s));
        }
        return VG_OK;
    }
    VG_REQUIRE(g.np[0] >= 2 && g.np[1] >= 2 && g.np[2] >= 2, "grid template is degenerate");
    VG_TRY(grow_f(&sim->vel_grid, &sim->vel_cap, 3 * cap));
    if (sim->cfg.technique == VG_TECH_POISSON) {
        VG_TRY(grow_f(&sim->a0, &sim->a0_cap, 3 * cap));
        VG_TRY(grow_f(&sim->negw, &sim->negw_cap, 3 * cap));
    }
    VG_TRY(vg_tree_build(sim->tree, &g, &sim->v, s));
    float radius = 0.f;
    VG_TRY(vorton_radius(sim, &radius));
    const int pol = sim->cfg.rsqrt_policy;
    if (sim->cfg.technique == VG_TECH_POISSON) {
        sim->spread = 1.f;
        sim->circ = 1.f;
        VG_TRY(vg_tree_splat_vorticity(sim->tree, radius, -1.0f, sim->negw, s));
        VG_TRY(zero_fill(sim->a0, 3 * cap * sizeof(float), s));
        VG_TRY(vg_tree_vecpot_boundary(sim->tree, &g, 1.f, 1.f, pol, sim->cfg.boundary_copy_odd,
                                       sim->a0, s));
        if (sim->cfg.poisson_solver == 0) {
            VG_TRY(vg_mg_cascade(sim->mg, &g, sim->a0, sim->negw, sim->cfg.mg_steps, s));
            sim->residual = -1.0;
        } else {
            uint32_t cycles = 0;
            VG_TRY(vg_mg_vcycle_solve(sim->mg, &g, sim->a0, sim->negw, sim->cfg.mg_tol,
                                      sim->cfg.mg_max_cycles, &sim->residual, &cycles, s));
        }
        VG_TRY(curl(g, sim->a0, sim->vel_grid, s));
    } else if (sim->cfg.technique == VG_TECH_TREE) {
        sim->spread = 1.f;
        sim->circ = 1.f;
        if (sim->cfg.auto_mollify) VG_TRY(vg_auto_mollify(&g, radius, &sim->spread, &sim->circ));
        if (sim->cfg.tree_near_k < 0)
            VG_TRY(vg_tree_velocity_grid(sim->tree, &g, sim->spread, sim->circ, pol, sim->vel_grid,
                                         s));
        else if (sim->mlp)
            VG_TRY(vg_tree_velocity_grid_mlp(sim->tree, &g, sim->mlp, sim->cfg.tree_near_k,
                                             sim->mlp_layers, sim->spread, sim->circ, pol,
                                             sim->vel_grid, s));
        else
            VG_TRY(vg_tree_velocity_grid_mac(sim->tree, &g, sim->cfg.tree_near_k,
                                             sim->cfg.tree_order, sim->cfg.tree_expand_from,
                                             sim->spread, sim->circ, pol,
                                             sim->vel_grid, s));
    } else {
        sim->spread = 1.f;
        sim->circ = 1.f;
        float *vv = vel_out;
        size_t vs = vel_stride;
        if (!vv) {
            VG_TRY(grow_f(&sim->vort_vel, &sim->vort_cap, 3 * sim->v.n));
            vv = sim->vort_vel;
            vs = 3;
        }
        if (vs == 3) {
            VG_TRY(vg_direct_velocity(&sim->v, sim->v.pos, sim->v.stride, sim->v.n, 1.f, 1.f, pol,
                                      vv, s));
        } else {
            // The direct kernel writes packed xyz; stage then scatter into the
            // caller's strided records.
            VG_TRY(grow_f(&sim->vort_vel, &sim->vort_cap, 3 * sim->v.n));
            VG_TRY(vg_direct_velocity(&sim->v, sim->v.pos, sim->v.stride, sim->v.n, 1.f, 1.f, pol,
                                      sim->vort_vel, s));
            VG_HIP(hipMemcpy2DAsync(vv, vs * sizeof(float), sim->vort_vel, 3 * sizeof(float),
                                    3 * sizeof(float), sim->v.n, hipMemcpyDeviceToDevice, s));
        }
        VG_TRY(vg_tree_populate_velocity_grid(sim->tree, vv, vs, sim->vel_grid, s));
        return VG_OK;
    }
    if (vel_out)
        VG_TRY(vg_assign_velocity_from_field(&g, sim->vel_grid, sim->v.pos, vel_out, vel_stride,
                                             sim->v.n, 1.0f, s));
    return VG_OK;
}

VG_API vg_status vg_sim_stretch_tilt(vg_sim *sim, float dt, float *angvel, size_t ang_stride) {
    VG_REQUIRE(sim && sim->have_vortons, "no vortons set");
    VG_REQUIRE(angvel && ang_stride >= 3, "bad angular velocity array");
    if (sim->v.n == 0 || capacity(sim->vel_desc) == 0) return VG_OK;
    VG_REQUIRE(sim->vel_grid, "no velocity grid: call vg_sim_update first");
    TreeView tv;
    VG_TRY(tree_checked_view(sim->tree, &tv));
    VG_REQUIRE(tv.n == sim->v.n, "vorton count changed since the update");
    return stretch_tilt_fused(sim->vel_desc, sim->vel_grid, sim->v.pos, sim->v.stride, angvel,
                              ang_stride, sim->v.n, dt, tv.sorted_idx, sim->stream);
}

VG_API vg_status vg_sim_pse_diffuse(vg_sim *sim, float viscosity, float dt, float *angvel,
                                    size_t ang_stride, const int32_t *birth,
                                    size_t birth_stride, int order) {
    VG_REQUIRE(sim && sim->have_vortons, "no vortons set");
    VG_REQUIRE(angvel && ang_stride >= 3, "bad angular velocity array");
    VG_REQUIRE(viscosity * dt <= 1.0f, "viscosity * dt exceeds 1 (vorticity would flip)");
    if (sim->v.n == 0) return VG_OK;
    TreeView tv;
    VG_TRY(tree_checked_view(sim->tree, &tv));
    VG_REQUIRE(tv.n == sim->v.n, "vorton count changed since the update");
    return pse_diffuse(sim->tree, angvel, ang_stride, birth, birth_stride, viscosity, dt, order,
                       sim->stream);
}

VG_API vg_status vg_sim_velocity_grid(const vg_sim *sim, vg_grid_desc *g, const float **field) {
    VG_REQUIRE(sim && g && field, "null argument");
    *g = sim->vel_desc;
    *field = sim->vel_grid;
    return VG_OK;
}

VG_API vg_status vg_sim_vector_potential(const vg_sim *sim, vg_grid_desc *g, const float **field) {
    VG_REQUIRE(sim && g && field, "null argument");
    *g = sim->tmpl;
    *field = sim->a0;
    return VG_OK;
}

VG_API vg_status vg_sim_multigrid(const vg_sim *sim, int field, uint32_t layer,
                                  vg_grid_desc *g, const float **data, uint32_t *depth) {
    VG_REQUIRE(sim && g && data, "null argument");
    VG_REQUIRE(field == VG_MG_VECTOR_POTENTIAL || field == VG_MG_NEGATIVE_VORTICITY,
               "field must be VG_MG_VECTOR_POTENTIAL or VG_MG_NEGATIVE_VORTICITY");
    uint32_t d = 0;
    if (sim->cfg.technique == VG_TECH_POISSON && sim->a0) {
        d = 1;
        if (sim->cfg.poisson_solver == 0 && mg_cascade_depth(sim->mg) > 1)
            d = mg_cascade_depth(sim->mg);
    }
    if (depth) *depth = d;
    VG_REQUIRE(layer < d, "layer beyond the pyramid the last update filled");
    if (layer == 0) {
        *g = sim->tmpl;
        *data = field == VG_MG_VECTOR_POTENTIAL ? sim->a0 : sim->negw;
        return VG_OK;
    }
    return vg_mg_layer(sim->mg, field, layer, g, data, nullptr);
}

VG_API vg_status vg_sim_negative_vorticity(const vg_sim *sim, vg_grid_desc *g,
                                           const float **field) {
    VG_REQUIRE(sim && g && field, "null argument");
    *g = sim->tmpl;
    *field = sim->negw;
    return VG_OK;
}

VG_API vg_status vg_sim_grid_template(const vg_sim *sim, vg_grid_desc *g) {
    VG_REQUIRE(sim && g, "null argument");
    *g = sim->tmpl;
    return VG_OK;
}

VG_API vg_status vg_sim_bounds(const vg_sim *sim, float mn[3], float mx[3]) {
    VG_REQUIRE(sim && mn && mx, "null argument");
    for (int a = 0; a < 3; ++a) { mn[a] = sim->mn[a]; mx[a] = sim->mx[a]; }
    return VG_OK;
}

VG_API vg_status vg_sim_tree(const vg_sim *sim, const vg_tree **t) {
    VG_REQUIRE(sim && t, "null argument");
    *t = sim->tree;
    return VG_OK;
}

VG_API vg_status vg_sim_spreading(const vg_sim *sim, float *spread, float *circ) {
    VG_REQUIRE(sim && spread && circ, "null argument");
    *spread = sim->spread;
    *circ = sim->circ;
    return VG_OK;
}

VG_API vg_status vg_sim_poisson_residual(const vg_sim *sim, double *rel) {
    VG_REQUIRE(sim && rel, "null argument");
    if (sim->residual >= 0.0) {
        *rel = sim->residual;
        return VG_OK;
    }
    if (!sim->a0 || !sim->negw || sim->cfg.technique != VG_TECH_POISSON) {
        *rel = -1.0;
        return VG_OK;
    }
    return residual_norm(sim->tmpl, sim->a0, sim->negw, rel, sim->stream);
}
