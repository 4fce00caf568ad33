/* vg_oracle.c -- CPU restatement of VorteGrid's velocity-from-vorticity path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  Compiled with
 * -ffp-contract=off so every float operation rounds as the reference source
 * text orders it (x86-64 SSE, FLT_EVAL_METHOD 0).  Citations are relative to
 * Samples/MjgIntelFluidDemo_Part20/.
 */
#include "vg_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

static const float kTwoThirds = 2.0f / 3.0f;           /* vorton.h:95 */
static const float kPi = 3.1415926535897932384626433832795f; /* particle.h:67 */

/* ------------------------------------------------------------------------ */
/* parallel-for over [0,n) in contiguous chunks (stands in for TBB's
 * parallel_for with grain n/threads; every item is written by one worker,
 * so results do not depend on the thread count). */
typedef void (*range_fn)(void *ctx, size_t b, size_t e);
typedef struct { range_fn fn; void *ctx; size_t b, e; } range_job;
static void *range_thread(void *a) {
    range_job *j = (range_job *)a;
    j->fn(j->ctx, j->b, j->e);
    return NULL;
}
static void parallel_range(size_t n, int nthreads, range_fn fn, void *ctx) {
    if (nthreads <= 1 || n < 2) { fn(ctx, 0, n); return; }
    if ((size_t)nthreads > n) nthreads = (int)n;
    pthread_t th[256];
    range_job jobs[256];
    if (nthreads > 256) nthreads = 256;
    size_t chunk = (n + nthreads - 1) / nthreads;
    int used = 0;
    for (int t = 0; t < nthreads; ++t) {
        size_t b = (size_t)t * chunk, e = b + chunk;
        if (b >= n) break;
        if (e > n) e = n;
        jobs[t].fn = fn; jobs[t].ctx = ctx; jobs[t].b = b; jobs[t].e = e;
        pthread_create(&th[t], NULL, range_thread, &jobs[t]);
        ++used;
    }
    for (int t = 0; t < used; ++t) pthread_join(th[t], NULL);
}

/* ------------------------------------------------------------------------ */
/* Core/Math/math.h:206-216.  The reference reads the float through `long&`,
 * which is 32-bit on its LLP64 target; restated with int32_t. */
float vgo_finvsqrtf(float val) {
    int32_t i;
    memcpy(&i, &val, 4);
    i = 0x5f3759df - (i >> 1);
    float y;
    memcpy(&y, &i, 4);
    const float f = 1.5f;
    const float x = val * 0.5f;
    y = y * (f - ((x * y) * y));
    return y;
}

/* Core/Math/math.h:149-175 */
uint32_t vgo_nearest_pow2(uint32_t v) {
    if (v == 0) return 1u;
    --v;
    unsigned shift = 0;
    while ((v >> shift) != 0) ++shift;
    return 1u << shift;
}

/* vorton.h:161-179 VORTON_ACCUMULATE_VELOCITY_private */
void vgo_accumulate_velocity(float acc[3], const float q[3], const float p[3],
                             const float om[3], float size, float spread,
                             float circ, int rsqrt_policy) {
    const float dx = q[0] - p[0], dy = q[1] - p[1], dz = q[2] - p[2];
    const float radius = size * 0.5f * spread;
    const float radius2 = radius * radius;
    const float dist2 = dx * dx + dy * dy + dz * dz;
    float law;
    if (dist2 < radius2) {
        law = 1.0f / (radius2 * radius);
    } else {
        const float rs = rsqrt_policy ? 1.0f / sqrtf(dist2) : vgo_finvsqrtf(dist2);
        law = rs / dist2;
    }
    /* (TwoThirds * radius2 * radius * Omega) ^ (d * law) */
    const float s = kTwoThirds * radius2 * radius;
    const float ax = om[0] * s, ay = om[1] * s, az = om[2] * s;
    const float bx = dx * law, by = dy * law, bz = dz * law;
    const float cx = ay * bz - az * by;
    const float cy = az * bx - ax * bz;
    const float cz = ax * by - ay * bx;
    acc[0] += cx * circ;
    acc[1] += cy * circ;
    acc[2] += cz * circ;
}

/* vorton.h:118-136 VORTON_ACCUMULATE_VECTOR_POTENTIAL_private */
void vgo_accumulate_vecpot(float acc[3], const float q[3], const float p[3],
                           const float om[3], float size, float spread,
                           float circ, int rsqrt_policy) {
    const float dx = q[0] - p[0], dy = q[1] - p[1], dz = q[2] - p[2];
    const float radius = size * 0.5f * spread;
    const float radius2 = radius * radius;
    const float dist2 = dx * dx + dy * dy + dz * dz;
    float law;
    if (dist2 < radius2) {
        law = dist2 / (radius2 * radius);
    } else {
        law = rsqrt_policy ? 1.0f / sqrtf(dist2) : vgo_finvsqrtf(dist2);
    }
    const float s = kTwoThirds * radius2 * radius;
    acc[0] += ((om[0] * s) * law) * circ;
    acc[1] += ((om[1] * s) * law) * circ;
    acc[2] += ((om[2] * s) * law) * circ;
}

/* ------------------------------------------------------------------------ */
/* vortonSim.cpp:1113-1158: serial sum over vortons 0..N-1 per query. */
typedef struct {
    const float *pos, *om, *size, *q;
    size_t n;
    float spread, circ;
    int policy, vecpot;
    float *out;
    double *out64;
} direct_ctx;

static void direct_range(void *vc, size_t b, size_t e) {
    direct_ctx *c = (direct_ctx *)vc;
    for (size_t iq = b; iq < e; ++iq) {
        float acc[3] = {0.f, 0.f, 0.f};
        const float *q = c->q + 3 * iq;
        for (size_t i = 0; i < c->n; ++i) {
            if (c->vecpot)
                vgo_accumulate_vecpot(acc, q, c->pos + 3 * i, c->om + 3 * i,
                                      c->size[i], c->spread, c->circ, c->policy);
            else
                vgo_accumulate_velocity(acc, q, c->pos + 3 * i, c->om + 3 * i,
                                        c->size[i], c->spread, c->circ, c->policy);
        }
        c->out[3 * iq] = acc[0];
        c->out[3 * iq + 1] = acc[1];
        c->out[3 * iq + 2] = acc[2];
    }
}

void vgo_direct_velocity(const float *pos, const float *om, const float *size,
                         size_t n, const float *q, size_t nq, float spread,
                         float circ, int rsqrt_policy, float *out, int nthreads) {
    direct_ctx c = {pos, om, size, q, n, spread, circ, rsqrt_policy, 0, out, NULL};
    parallel_range(nq, nthreads, direct_range, &c);
}

void vgo_direct_vecpot(const float *pos, const float *om, const float *size,
                       size_t n, const float *q, size_t nq, float spread,
                       float circ, int rsqrt_policy, float *out, int nthreads) {
    direct_ctx c = {pos, om, size, q, n, spread, circ, rsqrt_policy, 1, out, NULL};
    parallel_range(nq, nthreads, direct_range, &c);
}

/* FP64 accuracy oracle: the same regularised law as vorton.h:161-179 evaluated
 * in double with an exact reciprocal square root. */
static void direct64_range(void *vc, size_t b, size_t e) {
    direct_ctx *c = (direct_ctx *)vc;
    for (size_t iq = b; iq < e; ++iq) {
        double ax = 0, ay = 0, az = 0;
        const double qx = c->q[3 * iq], qy = c->q[3 * iq + 1], qz = c->q[3 * iq + 2];
        for (size_t i = 0; i < c->n; ++i) {
            const double dx = qx - c->pos[3 * i], dy = qy - c->pos[3 * i + 1],
                         dz = qz - c->pos[3 * i + 2];
            const double r = (double)c->size[i] * 0.5 * (double)c->spread;
            const double r2 = r * r;
            const double d2 = dx * dx + dy * dy + dz * dz;
            const double law = d2 < r2 ? 1.0 / (r2 * r) : 1.0 / (d2 * sqrt(d2));
            const double s = (2.0 / 3.0) * r2 * r * (double)c->circ;
            const double wx = c->om[3 * i] * s, wy = c->om[3 * i + 1] * s,
                         wz = c->om[3 * i + 2] * s;
            ax += (wy * dz - wz * dy) * law;
            ay += (wz * dx - wx * dz) * law;
            az += (wx * dy - wy * dx) * law;
        }
        c->out64[3 * iq] = ax;
        c->out64[3 * iq + 1] = ay;
        c->out64[3 * iq + 2] = az;
    }
}

void vgo_direct_velocity_f64(const float *pos, const float *om, const float *size,
                             size_t n, const float *q, size_t nq, float spread,
                             float circ, double *out, int nthreads) {
    direct_ctx c = {pos, om, size, q, n, spread, circ, 1, 0, NULL, out};
    parallel_range(nq, nthreads, direct64_range, &c);
}

/* ------------------------------------------------------------------------ */
/* uniformGrid.h:794-809 PrecomputeSpacing */
static void precompute_spacing(vgo_grid *g) {
    for (int a = 0; a < 3; ++a)
        g->cell[a] = g->extent[a] / (float)(g->np[a] - 1);
    g->cpe[0] = (float)(g->np[0] - 1) / g->extent[0];
    g->cpe[1] = (float)(g->np[1] - 1) / g->extent[1];
    if (0.0f == g->extent[2])
        g->cpe[2] = 1.0f / FLT_MIN;
    else
        g->cpe[2] = (float)(g->np[2] - 1) / g->extent[2];
}

/* The reference's powf calls of the grid geometry (DefineShape's cells per
 * extent, the auto-mollifier's cell size; uniformGrid.h:291,
 * vortonSim.cpp:2471) as pow in double rounded once to float: within half an
 * ulp of the exact power up to double rounding.  The reference's own result
 * is its platform powf's, which may differ from this in the last bit. */
static float pow_rounded(float x, float y) { return (float)pow((double)x, (double)y); }

/* uniformGrid.h:259-322 DefineShape */
void vgo_define_shape(vgo_grid *g, size_t n, const float vmin[3],
                      const float vmax[3], int pow2) {
    memset(g, 0, sizeof(*g));
    if (n == 0) return;
    const float nudge = 1.0f + FLT_EPSILON;
    float eff[3];
    int ndims = 3;
    for (int a = 0; a < 3; ++a) {
        g->min[a] = vmin[a];
        g->extent[a] = (vmax[a] - vmin[a]) * nudge;
        eff[a] = g->extent[a];
    }
    for (int a = 0; a < 3; ++a) {
        if (0.0f == eff[a]) { eff[a] = 1.0f; g->extent[a] = 0.0f; --ndims; }
    }
    const float volume = eff[0] * eff[1] * eff[2];
    const float icvcr = pow_rounded(volume / (float)n, -1.0f / (float)ndims);
    size_t nc[3];
    for (int a = 0; a < 3; ++a) {
        size_t v = (size_t)(g->extent[a] * icvcr + 0.5f);
        nc[a] = v > 1 ? v : 1;
    }
    if (pow2)
        for (int a = 0; a < 3; ++a) nc[a] = vgo_nearest_pow2((uint32_t)nc[a]);
    while (nc[0] * nc[1] * nc[2] >= n * 8) {
        for (int a = 0; a < 3; ++a) nc[a] = nc[a] / 2 > 1 ? nc[a] / 2 : 1;
    }
    for (int a = 0; a < 3; ++a) g->np[a] = (uint32_t)(nc[a] + 1);
    precompute_spacing(g);
}

/* uniformGrid.h:361-376 Decimate */
void vgo_decimate(vgo_grid *dst, const vgo_grid *src, uint32_t dec) {
    vgo_grid d;
    memset(&d, 0, sizeof(d));
    for (int a = 0; a < 3; ++a) {
        d.extent[a] = src->extent[a];
        d.min[a] = src->min[a];
        d.np[a] = (src->np[a] - 1) / dec + 1;
        if (dec > 1 && d.np[a] < 2) d.np[a] = 2;
    }
    precompute_spacing(&d);
    *dst = d;
}

size_t vgo_grid_capacity(const vgo_grid *g) {
    return (size_t)g->np[0] * g->np[1] * g->np[2];
}

void vgo_grid_query_point(const vgo_grid *g, uint32_t ix, uint32_t iy,
                          uint32_t iz, float out[3]) {
    const float nudge = 1.0f - 2.0f * FLT_EPSILON;
    const uint32_t idx[3] = {ix, iy, iz};
    for (int a = 0; a < 3; ++a)
        out[a] = g->min[a] + (float)idx[a] * (g->cell[a] * nudge);
}

/* ------------------------------------------------------------------------ */
/* pclOpFindBoundingBox.cpp:20-148 (serial form; min/max are exact). */
void vgo_find_bbox(const float *pos, size_t stride, size_t n, float mn[3],
                   float mx[3]) {
    for (size_t i = 0; i < n; ++i) {
        const float *p = pos + i * stride;
        for (int a = 0; a < 3; ++a) {
            if (p[a] < mn[a]) mn[a] = p[a];
            if (p[a] > mx[a]) mx[a] = p[a];
        }
    }
}

/* vortonSim.cpp:1027-1096 UpdateBoundingBox(..., bFinal=true) */
void vgo_finalize_bbox(float mn[3], float mx[3], float vorton_radius,
                       size_t n_vortons, vgo_grid *grid_template) {
    if (n_vortons == 0) {
        const float z[3] = {0.f, 0.f, 0.f};
        vgo_define_shape(grid_template, 0, z, z, 1);
        return;
    }
    const float margin = (1.0f + 2.0f) * vorton_radius;
    const float nudge = margin * 1.0f;
    for (int a = 0; a < 3; ++a) { mn[a] -= nudge; mx[a] += nudge; }
    vgo_define_shape(grid_template, n_vortons, mn, mx, 1);
}

/* ------------------------------------------------------------------------ */
/* nestedGrid.h:69-85 Initialize + PrecomputeDecimations (:371-391) */
void vgo_nested_init(vgo_nested *ng, const vgo_grid *src) {
    memset(ng, 0, sizeof(*ng));
    vgo_decimate(&ng->layer[0], src, 1);
    uint32_t idx = 1;
    while (vgo_grid_capacity(&ng->layer[idx - 1]) > 8 && idx < VGO_MAX_LAYERS) {
        vgo_decimate(&ng->layer[idx], &ng->layer[idx - 1], 2);
        ++idx;
    }
    ng->depth = idx;
    for (uint32_t l = 1; l < idx; ++l)
        for (int a = 0; a < 3; ++a)
            ng->dec[l][a] = (ng->layer[l - 1].np[a] - 1) / (ng->layer[l].np[a] - 1);
}

/* uniformGrid.h:536-574 IndicesOfPosition (truncation). */
static void indices_of_position(const vgo_grid *g, const float p[3],
                                uint32_t idx[3]) {
    for (int a = 0; a < 3; ++a) {
        const float rel = p[a] - g->min[a];
        const float f = rel * g->cpe[a];
        idx[a] = (uint32_t)(int32_t)f;
    }
}

void vgo_indices_of_position(const vgo_grid *g, const float p[3], uint32_t idx[3]) {
    indices_of_position(g, p, idx);
}

/* vortonSim.cpp:1192-1321 MakeBaseVortonGrid and :1337-1569 AggregateClusters */
int vgo_tree_build(vgo_tree *t, const vgo_grid *grid_template,
                   const float *pos, const float *om, const float *size,
                   size_t n) {
    memset(t, 0, sizeof(*t));
    vgo_nested_init(&t->ng, grid_template);
    for (uint32_t l = 0; l < t->ng.depth; ++l) {
        const size_t cap = vgo_grid_capacity(&t->ng.layer[l]);
        t->pos[l] = (float *)calloc(cap * 3, sizeof(float));
        t->om[l] = (float *)calloc(cap * 3, sizeof(float));
        t->size[l] = (float *)calloc(cap, sizeof(float));
        t->dip[l] = (float *)calloc(cap * 9, sizeof(float));
        if (!t->pos[l] || !t->om[l] || !t->size[l] || !t->dip[l]) return -1;
    }
    const vgo_grid *g0 = &t->ng.layer[0];
    const size_t cap0 = vgo_grid_capacity(g0);
    float *aux = (float *)malloc(cap0 * sizeof(float));
    for (size_t i = 0; i < cap0; ++i) aux[i] = FLT_MIN;
    for (size_t v = 0; v < n; ++v) {
        uint32_t idx[3];
        indices_of_position(g0, pos + 3 * v, idx);
        const size_t off = idx[0] + (size_t)g0->np[0] * (idx[1] + (size_t)g0->np[1] * idx[2]);
        const float wx = om[3 * v] * 2.0f, wy = om[3 * v + 1] * 2.0f, wz = om[3 * v + 2] * 2.0f;
        const float mag = sqrtf(wx * wx + wy * wy + wz * wz);
        for (int a = 0; a < 3; ++a) t->pos[0][3 * off + a] += pos[3 * v + a] * mag;
        for (int a = 0; a < 3; ++a) t->om[0][3 * off + a] += om[3 * v + a];
        t->size[0][off] = size[v];
        aux[off] += mag;
    }
    for (size_t off = 0; off < cap0; ++off) {
        if (aux[off] != FLT_MIN) {
            const float inv = 1.0f / aux[off];
            for (int a = 0; a < 3; ++a) t->pos[0][3 * off + a] *= inv;
        }
    }
    free(aux);
    /* Leaf cell lists (stable counting sort by cell) and leaf first moments
     * about the cell centroid, summed in vorton index order. */
    t->n = n;
    t->cell_start = (uint32_t *)calloc(cap0, sizeof(uint32_t));
    t->cell_count = (uint32_t *)calloc(cap0, sizeof(uint32_t));
    t->sorted_idx = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    uint32_t *keys = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    if (!t->cell_start || !t->cell_count || !t->sorted_idx || !keys) return -1;
    for (size_t v = 0; v < n; ++v) {
        uint32_t idx[3];
        indices_of_position(g0, pos + 3 * v, idx);
        keys[v] = (uint32_t)(idx[0] + (size_t)g0->np[0] * (idx[1] + (size_t)g0->np[1] * idx[2]));
        t->cell_count[keys[v]]++;
    }
    {
        uint32_t run = 0;
        for (size_t c = 0; c < cap0; ++c) { t->cell_start[c] = run; run += t->cell_count[c]; }
        uint32_t *fill = (uint32_t *)calloc(cap0, sizeof(uint32_t));
        for (size_t v = 0; v < n; ++v) {
            const uint32_t k = keys[v];
            t->sorted_idx[t->cell_start[k] + fill[k]++] = (uint32_t)v;
        }
        free(fill);
    }
    free(keys);
    for (size_t c = 0; c < cap0; ++c) {
        const float *cp = t->pos[0] + 3 * c;
        float *d = t->dip[0] + 9 * c;
        for (uint32_t k = 0; k < t->cell_count[c]; ++k) {
            const size_t v = t->sorted_idx[t->cell_start[c] + k];
            const float e[3] = {pos[3 * v] - cp[0], pos[3 * v + 1] - cp[1], pos[3 * v + 2] - cp[2]};
            for (int b = 0; b < 3; ++b)
                for (int a = 0; a < 3; ++a) d[3 * b + a] += om[3 * v + b] * e[a];
        }
    }
    for (uint32_t pl = 1; pl < t->ng.depth; ++pl) {
        const vgo_grid *gp = &t->ng.layer[pl], *gc = &t->ng.layer[pl - 1];
        const uint32_t *dims = t->ng.dec[pl];
        const uint32_t nc[3] = {gp->np[0] - 1, gp->np[1] - 1, gp->np[2] - 1};
        for (uint32_t iz = 0; iz < nc[2]; ++iz)
            for (uint32_t iy = 0; iy < nc[1]; ++iy)
                for (uint32_t ix = 0; ix < nc[0]; ++ix) {
                    const size_t po = ix + (size_t)gp->np[0] * (iy + (size_t)gp->np[1] * iz);
                    float *pp = t->pos[pl] + 3 * po, *pw = t->om[pl] + 3 * po;
                    float asum = FLT_MIN;
                    for (uint32_t cz = 0; cz < dims[2]; ++cz)
                        for (uint32_t cy = 0; cy < dims[1]; ++cy)
                            for (uint32_t cx = 0; cx < dims[0]; ++cx) {
                                const size_t co = (ix * dims[0] + cx) +
                                    (size_t)gc->np[0] * ((iy * dims[1] + cy) +
                                    (size_t)gc->np[1] * (iz * dims[2] + cz));
                                const float *cw = t->om[pl - 1] + 3 * co;
                                const float wx = cw[0] * 2.0f, wy = cw[1] * 2.0f, wz = cw[2] * 2.0f;
                                const float mag = sqrtf(wx * wx + wy * wy + wz * wz);
                                for (int a = 0; a < 3; ++a)
                                    pp[a] += t->pos[pl - 1][3 * co + a] * mag;
                                for (int a = 0; a < 3; ++a) pw[a] += cw[a];
                                asum += mag;
                                if (t->size[pl - 1][co] != 0.0f)
                                    t->size[pl][po] = t->size[pl - 1][co];
                            }
                    const float inv = 1.0f / asum;
                    for (int a = 0; a < 3; ++a) pp[a] *= inv;
                    /* first moment about the new centroid (parallel-axis shift) */
                    float *pd = t->dip[pl] + 9 * po;
                    for (uint32_t cz = 0; cz < dims[2]; ++cz)
                        for (uint32_t cy = 0; cy < dims[1]; ++cy)
                            for (uint32_t cx = 0; cx < dims[0]; ++cx) {
                                const size_t co = (ix * dims[0] + cx) +
                                    (size_t)gc->np[0] * ((iy * dims[1] + cy) +
                                    (size_t)gc->np[1] * (iz * dims[2] + cz));
                                const float *cw = t->om[pl - 1] + 3 * co;
                                const float *cpos = t->pos[pl - 1] + 3 * co;
                                const float *cd = t->dip[pl - 1] + 9 * co;
                                for (int b = 0; b < 3; ++b)
                                    for (int a = 0; a < 3; ++a)
                                        pd[3 * b + a] += cd[3 * b + a] + cw[b] * (cpos[a] - pp[a]);
                            }
                }
    }
    return 0;
}

void vgo_tree_free(vgo_tree *t) {
    for (uint32_t l = 0; l < VGO_MAX_LAYERS; ++l) {
        free(t->pos[l]); free(t->om[l]); free(t->size[l]); free(t->dip[l]);
        t->pos[l] = t->om[l] = t->size[l] = t->dip[l] = NULL;
    }
    free(t->cell_start); free(t->cell_count); free(t->sorted_idx);
    t->cell_start = t->cell_count = t->sorted_idx = NULL;
}

/* ------------------------------------------------------------------------ */
/* Near-field treecode over 4x4x4 grid-point bricks (accurate mode). */
static void kahan3(float sum[3], float comp[3], const float part[3]) {
    for (int a = 0; a < 3; ++a) {
        const float y = part[a] - comp[a];
        const float t = sum[a] + y;
        comp[a] = (t - sum[a]) - y;
        sum[a] = t;
    }
}

static void cluster_contrib(float acc[3], const float q[3], const float *cp, const float *w,
                            float size, const float *D, int order, float spread, int policy) {
    const float radius = size * 0.5f * spread;
    const float r2 = radius * radius;
    const float s = (2.0f / 3.0f) * r2 * radius;
    const float d[3] = {q[0] - cp[0], q[1] - cp[1], q[2] - cp[2]};
    const float d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    float law;
    if (d2 < r2) {
        law = 1.0f / (r2 * radius);
    } else if (policy) {
        const float rs = 1.0f / sqrtf(d2);
        law = rs * rs * rs;
    } else {
        law = vgo_finvsqrtf(d2) / d2;
    }
    const float W[3] = {w[0] * s, w[1] * s, w[2] * s};
    acc[0] += (W[1] * d[2] - W[2] * d[1]) * law;
    acc[1] += (W[2] * d[0] - W[0] * d[2]) * law;
    acc[2] += (W[0] * d[1] - W[1] * d[0]) * law;
    if (order >= 1 && d2 >= r2) {
        /* u -= s/|d|^3 [ a - 3/|d|^2 (D d) x d ],  a = antisymmetric part of D */
        const float a[3] = {D[5] - D[7], D[6] - D[2], D[1] - D[3]};
        const float tt[3] = {D[0] * d[0] + D[1] * d[1] + D[2] * d[2],
                             D[3] * d[0] + D[4] * d[1] + D[5] * d[2],
                             D[6] * d[0] + D[7] * d[1] + D[8] * d[2]};
        const float c3 = 3.0f / d2;
        const float x[3] = {tt[1] * d[2] - tt[2] * d[1], tt[2] * d[0] - tt[0] * d[2],
                            tt[0] * d[1] - tt[1] * d[0]};
        const float sl = s * law;
        for (int k = 0; k < 3; ++k) acc[k] -= (a[k] - c3 * x[k]) * sl;
    }
}

typedef struct {
    const vgo_tree *t;
    const vgo_grid *g;
    const float *pos, *om, *size;
    int k, order, policy, expand_from;
    float spread, circ;
    float *out;
    uint32_t nb[3];
    uint32_t bx0, bx1; /* brick columns evaluated (timing samples) */
} mac_ctx;

/* Local expansion about a brick centre (no reference counterpart), in
 * float64 from the tensors of K(d) = d/|d|^3:
 *   J_ij  = d_j K_i  = delta_ij/r^3 - 3 d_i d_j/r^5
 *   T_ijk = d_k J_ij = -3 (delta_ij d_k + delta_ik d_j + delta_jk d_i)/r^5
 *                      + 15 d_i d_j d_k/r^7
 * u(x) = s [Om x K(d) - sum_be D_be e_b x d_e K(d)], d = x - c: the monopole
 * term is kept to second order in x - c_b, the first-moment term to first
 * order.  c0[i], c1[i][j] = du_i/dx_j, c2[i][j][k] = d2u_i/dx_j dx_k. */
typedef struct { double c0[3], c1[3][3], c2[3][3][3]; } expansion;

static void cross_d(const double a[3], const double b[3], double o[3]) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

static void expansion_add(expansion *E, const float cb[3], const float *cp, const float *w,
                          float size, const float *D, int order, float spread) {
    double d[3], J[3][3], T[3][3][3];
    for (int a = 0; a < 3; ++a) d[a] = (double)cb[a] - (double)cp[a];
    const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    const double r = sqrt(r2), ir3 = 1.0 / (r2 * r), ir5 = ir3 / r2, ir7 = ir5 / r2;
    const double radius = (double)size * 0.5 * (double)spread;
    const double s = (2.0 / 3.0) * radius * radius * radius;
    double K[3];
    for (int i = 0; i < 3; ++i) {
        K[i] = d[i] * ir3;
        for (int j = 0; j < 3; ++j) {
            J[i][j] = (i == j ? ir3 : 0.0) - 3.0 * d[i] * d[j] * ir5;
            for (int k = 0; k < 3; ++k)
                T[i][j][k] = -3.0 * ((i == j ? d[k] : 0.0) + (i == k ? d[j] : 0.0) +
                                     (j == k ? d[i] : 0.0)) * ir5 +
                             15.0 * d[i] * d[j] * d[k] * ir7;
        }
    }
    const double om[3] = {s * w[0], s * w[1], s * w[2]};
    double v[3], o[3];
    cross_d(om, K, o);
    for (int i = 0; i < 3; ++i) E->c0[i] += o[i];
    for (int j = 0; j < 3; ++j) {
        for (int i = 0; i < 3; ++i) v[i] = J[i][j];
        cross_d(om, v, o);
        for (int i = 0; i < 3; ++i) E->c1[i][j] += o[i];
        for (int k = 0; k < 3; ++k) {
            for (int i = 0; i < 3; ++i) v[i] = T[i][j][k];
            cross_d(om, v, o);
            for (int i = 0; i < 3; ++i) E->c2[i][j][k] += o[i];
        }
    }
    if (order >= 1) {
        /* -s sum_b e_b x (sum_e D_be dK/dx_e) and its gradient */
        for (int b = 0; b < 3; ++b) {
            double eb[3] = {0.0, 0.0, 0.0};
            eb[b] = 1.0;
            double g0[3] = {0.0, 0.0, 0.0};
            for (int e = 0; e < 3; ++e)
                for (int i = 0; i < 3; ++i) g0[i] += s * (double)D[3 * b + e] * J[i][e];
            cross_d(eb, g0, o);
            for (int i = 0; i < 3; ++i) E->c0[i] -= o[i];
            for (int j = 0; j < 3; ++j) {
                double g1[3] = {0.0, 0.0, 0.0};
                for (int e = 0; e < 3; ++e)
                    for (int i = 0; i < 3; ++i) g1[i] += s * (double)D[3 * b + e] * T[i][e][j];
                cross_d(eb, g1, o);
                for (int i = 0; i < 3; ++i) E->c1[i][j] -= o[i];
            }
        }
    }
}

void vgo_local_expansion(const float cb[3], const float cp[3], const float w[3], float size,
                         const float D[9], int order, float spread, double out[39]) {
    expansion E;
    memset(&E, 0, sizeof E);
    expansion_add(&E, cb, cp, w, size, D, order, spread);
    memcpy(out, &E, sizeof E);
}

static void expansion_eval(const expansion *E, const float cb[3], const float q[3], float out[3]) {
    double x[3];
    for (int a = 0; a < 3; ++a) x[a] = (double)q[a] - (double)cb[a];
    for (int i = 0; i < 3; ++i) {
        double u = E->c0[i];
        for (int j = 0; j < 3; ++j) {
            u += E->c1[i][j] * x[j];
            for (int k = 0; k < 3; ++k) u += 0.5 * E->c2[i][j][k] * x[j] * x[k];
        }
        out[i] = (float)u;
    }
}

static void mac_brick(mac_ctx *c, uint32_t bx, uint32_t by, uint32_t bz) {
    const vgo_tree *t = c->t;
    const vgo_grid *g = c->g;
    const uint32_t depth = t->ng.depth;
    const uint32_t b0[3] = {4 * bx, 4 * by, 4 * bz};
    uint32_t b1[3];
    const uint32_t bb[3] = {bx, by, bz};
    for (int a = 0; a < 3; ++a) {
        b1[a] = 4 * bb[a] + 3;
        if (b1[a] > g->np[a] - 1) b1[a] = g->np[a] - 1;
    }
    /* leaf cells of the brick's first and last points; cumulative scales */
    int cl0[3], cl1[3];
    for (int a = 0; a < 3; ++a) {
        const int nc = (int)g->np[a] - 1;
        cl0[a] = (int)b0[a] < nc - 1 ? (int)b0[a] : nc - 1;
        cl1[a] = (int)b1[a] < nc - 1 ? (int)b1[a] : nc - 1;
    }
    uint32_t scale[VGO_MAX_LAYERS][3];
    for (int a = 0; a < 3; ++a) scale[0][a] = 1;
    for (uint32_t l = 1; l < depth; ++l)
        for (int a = 0; a < 3; ++a) scale[l][a] = scale[l - 1][a] * t->ng.dec[l][a];
    /* brick-level local expansion of the layers >= expand_from */
    const int expanding = c->expand_from >= 1 && c->expand_from <= (int)depth - 2;
    expansion E;
    memset(&E, 0, sizeof E);
    float cb[3];
    {
        const float nudge = 1.0f - 2.0f * FLT_EPSILON;
        for (int a = 0; a < 3; ++a)
            cb[a] = g->min[a] + ((float)(4 * bb[a]) + 1.5f) * (g->cell[a] * nudge);
    }
    if (expanding) {
        int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
        for (int l = (int)depth - 2; l >= c->expand_from; --l) {
            const vgo_grid *gl = &t->ng.layer[l];
            const uint32_t *dec = t->ng.dec[l + 1];
            int c0[3], c1[3], n0[3], n1[3];
            for (int a = 0; a < 3; ++a) {
                c0[a] = lo[a] * (int)dec[a];
                c1[a] = hi[a] * (int)dec[a] + (int)dec[a] - 1;
                const int ncl = (int)gl->np[a] - 1;
                int blo = cl0[a] / (int)scale[l][a] - c->k;
                int bhi = cl1[a] / (int)scale[l][a] + c->k;
                if (blo < 0) blo = 0;
                if (bhi > ncl - 1) bhi = ncl - 1;
                n0[a] = blo > c0[a] ? blo : c0[a];
                n1[a] = bhi < c1[a] ? bhi : c1[a];
            }
            for (int cz = c0[2]; cz <= c1[2]; ++cz)
                for (int cy = c0[1]; cy <= c1[1]; ++cy)
                    for (int cx = c0[0]; cx <= c1[0]; ++cx) {
                        if (cx >= n0[0] && cx <= n1[0] && cy >= n0[1] && cy <= n1[1] &&
                            cz >= n0[2] && cz <= n1[2])
                            continue;
                        const size_t off = cx + (size_t)gl->np[0] * (cy + (size_t)gl->np[1] * cz);
                        const float sz = t->size[l][off];
                        if (sz == 0.0f) continue;
                        expansion_add(&E, cb, t->pos[l] + 3 * off, t->om[l] + 3 * off, sz,
                                      t->dip[l] + 9 * off, c->order, c->spread);
                    }
            for (int a = 0; a < 3; ++a) { lo[a] = n0[a]; hi[a] = n1[a]; }
        }
    }
    for (uint32_t pz = b0[2]; pz <= b1[2]; ++pz)
        for (uint32_t py = b0[1]; py <= b1[1]; ++py)
            for (uint32_t px = b0[0]; px <= b1[0]; ++px) {
                float q[3];
                vgo_grid_query_point(g, px, py, pz, q);
                float sum[3] = {0, 0, 0}, comp[3] = {0, 0, 0};
                int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
                if (depth == 1) {
                    float part[3] = {0, 0, 0};
                    for (size_t v = 0; v < t->n; ++v)
                        vgo_accumulate_velocity(part, q, c->pos + 3 * v, c->om + 3 * v, c->size[v],
                                                c->spread, 1.0f, c->policy);
                    kahan3(sum, comp, part);
                }
                for (int l = (int)depth - 2; l >= 0; --l) {
                    const vgo_grid *gl = &t->ng.layer[l];
                    const uint32_t *dec = t->ng.dec[l + 1];
                    int c0[3], c1[3], n0[3], n1[3];
                    for (int a = 0; a < 3; ++a) {
                        c0[a] = lo[a] * (int)dec[a];
                        c1[a] = hi[a] * (int)dec[a] + (int)dec[a] - 1;
                        const int ncl = (int)gl->np[a] - 1;
                        int blo = cl0[a] / (int)scale[l][a] - c->k;
                        int bhi = cl1[a] / (int)scale[l][a] + c->k;
                        if (blo < 0) blo = 0;
                        if (bhi > ncl - 1) bhi = ncl - 1;
                        n0[a] = blo > c0[a] ? blo : c0[a];
                        n1[a] = bhi < c1[a] ? bhi : c1[a];
                    }
                    if (expanding && l >= c->expand_from) {
                        for (int a = 0; a < 3; ++a) { lo[a] = n0[a]; hi[a] = n1[a]; }
                        if (l == c->expand_from) {
                            float ue[3];
                            expansion_eval(&E, cb, q, ue);
                            kahan3(sum, comp, ue);
                        }
                        continue;
                    }
                    float part[3] = {0, 0, 0};
                    for (int cz = c0[2]; cz <= c1[2]; ++cz)
                        for (int cy = c0[1]; cy <= c1[1]; ++cy)
                            for (int cx = c0[0]; cx <= c1[0]; ++cx) {
                                if (cx >= n0[0] && cx <= n1[0] && cy >= n0[1] && cy <= n1[1] &&
                                    cz >= n0[2] && cz <= n1[2])
                                    continue;
                                const size_t off = cx + (size_t)gl->np[0] * (cy + (size_t)gl->np[1] * cz);
                                const float sz = t->size[l][off];
                                if (sz == 0.0f) continue;
                                cluster_contrib(part, q, t->pos[l] + 3 * off, t->om[l] + 3 * off, sz,
                                                t->dip[l] + 9 * off, c->order, c->spread, c->policy);
                            }
                    kahan3(sum, comp, part);
                    for (int a = 0; a < 3; ++a) { lo[a] = n0[a]; hi[a] = n1[a]; }
                    if (l == 0) {
                        float pp[3] = {0, 0, 0};
                        for (int cz = n0[2]; cz <= n1[2]; ++cz)
                            for (int cy = n0[1]; cy <= n1[1]; ++cy)
                                for (int cx = n0[0]; cx <= n1[0]; ++cx) {
                                    const size_t off = cx + (size_t)gl->np[0] * (cy + (size_t)gl->np[1] * cz);
                                    for (uint32_t j = 0; j < t->cell_count[off]; ++j) {
                                        const size_t v = t->sorted_idx[t->cell_start[off] + j];
                                        vgo_accumulate_velocity(pp, q, c->pos + 3 * v, c->om + 3 * v,
                                                                c->size[v], c->spread, 1.0f,
                                                                c->policy);
                                    }
                                }
                        kahan3(sum, comp, pp);
                    }
                }
                const size_t o = px + (size_t)g->np[0] * (py + (size_t)g->np[1] * pz);
                for (int a = 0; a < 3; ++a) c->out[3 * o + a] = sum[a] * c->circ;
            }
}

typedef struct { mac_ctx *c; uint32_t bz0; } mac_slab;

/* Work units are single bricks, so one z row of bricks still spreads over
 * every thread. */
static void mac_slab_range(void *vs, size_t b, size_t e) {
    mac_slab *s = (mac_slab *)vs;
    mac_ctx *c = s->c;
    const size_t nbx = c->bx1 - c->bx0;
    for (size_t u = b; u < e; ++u) {
        const uint32_t bx = c->bx0 + (uint32_t)(u % nbx);
        const uint32_t by = (uint32_t)((u / nbx) % c->nb[1]);
        const uint32_t bz = s->bz0 + (uint32_t)(u / (nbx * c->nb[1]));
        mac_brick(c, bx, by, bz);
    }
}

int vgo_nested_has_orphans(const vgo_nested *ng) {
    for (uint32_t l = 1; l < ng->depth; ++l)
        for (int a = 0; a < 3; ++a)
            if (ng->layer[l - 1].np[a] - 1 != ng->dec[l][a] * (ng->layer[l].np[a] - 1)) return 1;
    return 0;
}

int vgo_mac_velocity_grid_box(const vgo_tree *t, const vgo_grid *g, const float *pos,
                              const float *om, const float *size, int near_k, int order,
                              int expand_from, float spread, float circ, int rsqrt_policy,
                              float *out, int nthreads, uint32_t bx_begin, uint32_t bx_end,
                              uint32_t bz_begin, uint32_t bz_end) {
    if (vgo_nested_has_orphans(&t->ng)) return -1;
    mac_ctx c;
    c.t = t; c.g = g; c.pos = pos; c.om = om; c.size = size;
    c.k = near_k; c.order = order; c.policy = rsqrt_policy; c.expand_from = expand_from;
    c.spread = spread; c.circ = circ; c.out = out;
    for (int a = 0; a < 3; ++a) c.nb[a] = (g->np[a] + 3) / 4;
    c.bx0 = bx_begin;
    c.bx1 = bx_end < c.nb[0] ? bx_end : c.nb[0];
    if (bz_end > c.nb[2]) bz_end = c.nb[2];
    if (bz_begin >= bz_end) return 0;
    if (c.bx0 >= c.bx1) return 0;
    mac_slab s = {&c, bz_begin};
    parallel_range((size_t)(bz_end - bz_begin) * c.nb[1] * (c.bx1 - c.bx0), nthreads,
                   mac_slab_range, &s);
    return 0;
}

int vgo_mac_velocity_grid_ex(const vgo_tree *t, const vgo_grid *g, const float *pos,
                             const float *om, const float *size, int near_k, int order,
                             int expand_from, float spread, float circ, int rsqrt_policy,
                             float *out, int nthreads, uint32_t bz_begin, uint32_t bz_end) {
    return vgo_mac_velocity_grid_box(t, g, pos, om, size, near_k, order, expand_from, spread,
                                     circ, rsqrt_policy, out, nthreads, 0, UINT32_MAX, bz_begin,
                                     bz_end);
}

int vgo_mac_velocity_grid_bricks(const vgo_tree *t, const vgo_grid *g, const float *pos,
                                 const float *om, const float *size, int near_k, int order,
                                 float spread, float circ, int rsqrt_policy, float *out,
                                 int nthreads, uint32_t bz_begin, uint32_t bz_end) {
    return vgo_mac_velocity_grid_ex(t, g, pos, om, size, near_k, order, -1, spread, circ,
                                    rsqrt_policy, out, nthreads, bz_begin, bz_end);
}

int vgo_mac_velocity_grid(const vgo_tree *t, const vgo_grid *g, const float *pos,
                          const float *om, const float *size, int near_k, int order,
                          float spread, float circ, int rsqrt_policy, float *out,
                          int nthreads) {
    return vgo_mac_velocity_grid_bricks(t, g, pos, om, size, near_k, order, spread, circ,
                                        rsqrt_policy, out, nthreads, 0, (g->np[2] + 3) / 4);
}

/* vortonSim.cpp:1650-1768 / 1787-1905: recursive traversal.  A cluster of
 * size 0 is empty (zero circulation); it is skipped, which leaves the sum
 * unchanged except that a query exactly at the origin no longer produces
 * 0/0 through the fast reciprocal square root. */
static void tree_rec(const vgo_tree *t, const float q[3], const uint32_t ind[3],
                     uint32_t ilayer, float spread, float circ, int policy,
                     int vecpot, float out[3]) {
    const vgo_grid *gc = &t->ng.layer[ilayer - 1];
    const uint32_t *dims = t->ng.dec[ilayer];
    const uint32_t cmin[3] = {ind[0] * dims[0], ind[1] * dims[1], ind[2] * dims[2]};
    const float mf = 0.0001f;
    float margin[3];
    for (int a = 0; a < 3; ++a) margin[a] = mf * gc->cell[a];
    if (0.0f == gc->cell[2]) margin[2] += FLT_MIN;
    float acc[3] = {0.f, 0.f, 0.f};
    for (uint32_t iz = 0; iz < dims[2]; ++iz) {
        const uint32_t cz = cmin[2] + iz;
        const float zmin = gc->min[2] + (float)cz * gc->cell[2];
        const float zmax = gc->min[2] + (float)(cz + 1) * gc->cell[2];
        for (uint32_t iy = 0; iy < dims[1]; ++iy) {
            const uint32_t cy = cmin[1] + iy;
            const float ymin = gc->min[1] + (float)cy * gc->cell[1];
            const float ymax = gc->min[1] + (float)(cy + 1) * gc->cell[1];
            for (uint32_t ix = 0; ix < dims[0]; ++ix) {
                const uint32_t cx = cmin[0] + ix;
                const float xmin = gc->min[0] + (float)cx * gc->cell[0];
                const float xmax = gc->min[0] + (float)(cx + 1) * gc->cell[0];
                if (ilayer > 1 && q[0] >= xmin - margin[0] && q[1] >= ymin - margin[1] &&
                    q[2] >= zmin - margin[2] && q[0] < xmax + margin[0] &&
                    q[1] < ymax + margin[1] && q[2] < zmax + margin[2]) {
                    const uint32_t ci[3] = {cx, cy, cz};
                    float sub[3];
                    tree_rec(t, q, ci, ilayer - 1, spread, circ, policy, vecpot, sub);
                    acc[0] += sub[0]; acc[1] += sub[1]; acc[2] += sub[2];
                } else {
                    const size_t off = cx + (size_t)gc->np[0] * (cy + (size_t)gc->np[1] * cz);
                    const float sz = t->size[ilayer - 1][off];
                    if (sz == 0.0f) continue;
                    if (vecpot)
                        vgo_accumulate_vecpot(acc, q, t->pos[ilayer - 1] + 3 * off,
                                              t->om[ilayer - 1] + 3 * off, sz, spread,
                                              circ, policy);
                    else
                        vgo_accumulate_velocity(acc, q, t->pos[ilayer - 1] + 3 * off,
                                                t->om[ilayer - 1] + 3 * off, sz, spread,
                                                circ, policy);
                }
            }
        }
    }
    out[0] = acc[0]; out[1] = acc[1]; out[2] = acc[2];
}

void vgo_tree_velocity(const vgo_tree *t, const float q[3], float spread,
                       float circ, int rsqrt_policy, float out[3]) {
    const uint32_t zeros[3] = {0, 0, 0};
    tree_rec(t, q, zeros, t->ng.depth - 1, spread, circ, rsqrt_policy, 0, out);
}

void vgo_tree_vecpot(const vgo_tree *t, const float q[3], float spread,
                     float circ, int rsqrt_policy, float out[3]) {
    const uint32_t zeros[3] = {0, 0, 0};
    tree_rec(t, q, zeros, t->ng.depth - 1, spread, circ, rsqrt_policy, 1, out);
}

/* vortonSim.cpp:2458-2477 */
void vgo_auto_mollify(const vgo_grid *g, float vorton_radius, float *spread,
                      float *circ) {
    const float one_third = 1.0f / 3.0f;
    const float cell_volume = g->cell[0] * g->cell[1] * g->cell[2];
    const float cgl = 1.0f * pow_rounded(cell_volume, one_third);
    float s = cgl / vorton_radius;
    if (s < 1.0f) s = 1.0f;
    *spread = s;
    *circ = 1.0f / (s * s * s);
}

/* ------------------------------------------------------------------------ */
typedef struct {
    const vgo_tree *t;
    const vgo_grid *g;
    const float *pos, *om, *size;
    size_t n;
    int technique, policy, copy_odd;
    float spread, circ;
    float *out;
} grid_ctx;

static void velgrid_range(void *vc, size_t b, size_t e) {
    grid_ctx *c = (grid_ctx *)vc;
    const vgo_grid *g = c->g;
    for (size_t iz = b; iz < e; ++iz)
        for (uint32_t iy = 0; iy < g->np[1]; ++iy)
            for (uint32_t ix = 0; ix < g->np[0]; ++ix) {
                float q[3], v[3] = {0.f, 0.f, 0.f};
                vgo_grid_query_point(g, ix, iy, (uint32_t)iz, q);
                const size_t off = ix + (size_t)g->np[0] * (iy + (size_t)g->np[1] * iz);
                if (c->technique == 1 && c->t && c->t->ng.depth > 1) {
                    vgo_tree_velocity(c->t, q, c->spread, c->circ, c->policy, v);
                } else {
                    for (size_t i = 0; i < c->n; ++i)
                        vgo_accumulate_velocity(v, q, c->pos + 3 * i, c->om + 3 * i,
                                                c->size[i], c->spread, c->circ, c->policy);
                }
                c->out[3 * off] = v[0];
                c->out[3 * off + 1] = v[1];
                c->out[3 * off + 2] = v[2];
            }
}

void vgo_velocity_grid(const vgo_tree *t, const vgo_grid *g, const float *pos,
                       const float *om, const float *size, size_t n,
                       int technique, float spread, float circ,
                       int rsqrt_policy, float *out, int nthreads) {
    grid_ctx c = {t, g, pos, om, size, n, technique, rsqrt_policy, 0, spread, circ, out};
    parallel_range(g->np[2], nthreads, velgrid_range, &c);
}

/* vortonSim.cpp:2055-2138 (boundariesOnly) */
static void vpb_point(grid_ctx *c, uint32_t ix, uint32_t iy, uint32_t iz) {
    const vgo_grid *g = c->g;
    const size_t off = ix + (size_t)g->np[0] * (iy + (size_t)g->np[1] * iz);
    if (c->copy_odd && (ix & 1u)) {
        for (int a = 0; a < 3; ++a) c->out[3 * off + a] = c->out[3 * (off - 1) + a];
        return;
    }
    float q[3], v[3] = {0.f, 0.f, 0.f};
    vgo_grid_query_point(g, ix, iy, iz, q);
    if (c->t && c->t->ng.depth > 1) {
        vgo_tree_vecpot(c->t, q, c->spread, c->circ, c->policy, v);
    } else {
        for (size_t i = 0; i < c->n; ++i)
            vgo_accumulate_vecpot(v, q, c->pos + 3 * i, c->om + 3 * i, c->size[i],
                                  c->spread, c->circ, c->policy);
    }
    for (int a = 0; a < 3; ++a) c->out[3 * off + a] = v[a];
}

static void vpb_range(void *vc, size_t b, size_t e) {
    grid_ctx *c = (grid_ctx *)vc;
    const vgo_grid *g = c->g;
    const uint32_t *d = g->np;
    for (size_t izs = b; izs < e; ++izs) {
        const uint32_t iz = (uint32_t)izs;
        const int tb = (iz == 0) || (iz == d[2] - 1);
        for (uint32_t iy = 0; iy < d[1]; ++iy) {
            const int fb = (iy == 0) || (iy == d[1] - 1);
            const uint32_t inc = (tb || fb) ? 1u : (d[0] - 1);
            for (uint32_t ix = 0; ix < d[0]; ix += inc) {
                vpb_point(c, ix, iy, iz);
                if (inc == 0) break;
            }
        }
    }
}

void vgo_vecpot_boundary(const vgo_tree *t, const vgo_grid *g,
                         const float *pos, const float *om, const float *size,
                         size_t n, float spread, float circ, int rsqrt_policy,
                         int copy_odd, float *out, int nthreads) {
    grid_ctx c = {t, g, pos, om, size, n, 1, rsqrt_policy, copy_odd, spread, circ, out};
    parallel_range(g->np[2], nthreads, vpb_range, &c);
}

/* ------------------------------------------------------------------------ */
/* uniformGrid.h:1117-1148 Interpolate (weights not clamped). */
void vgo_interpolate(const vgo_grid *g, const float *f, const float p[3],
                     float out[3]) {
    uint32_t idx[3];
    indices_of_position(g, p, idx);
    float t[3], o[3];
    for (int a = 0; a < 3; ++a) {
        const float cmin = g->min[a] + (float)idx[a] * g->cell[a];
        const float diff = p[a] - cmin;
        t[a] = diff * g->cpe[a];
        o[a] = 1.0f - t[a];
    }
    const size_t nx = g->np[0], nxy = (size_t)g->np[0] * g->np[1];
    const size_t o000 = idx[0] + nx * (idx[1] + (size_t)g->np[1] * idx[2]);
    const size_t o100 = o000 + 1, o010 = o000 + nx, o110 = o000 + nx + 1;
    const size_t o001 = o000 + nxy, o101 = o000 + nxy + 1, o011 = o000 + nxy + nx,
                 o111 = o000 + nxy + nx + 1;
    for (int c = 0; c < 3; ++c) {
        out[c] = ((o[0] * f[3 * o000 + c] + t[0] * f[3 * o100 + c]) * o[1] +
                  (o[0] * f[3 * o010 + c] + t[0] * f[3 * o110 + c]) * t[1]) * o[2] +
                 ((o[0] * f[3 * o001 + c] + t[0] * f[3 * o101 + c]) * o[1] +
                  (o[0] * f[3 * o011 + c] + t[0] * f[3 * o111 + c]) * t[1]) * t[2];
    }
}

static float clamp01(float v) { return v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v); }

/* uniformGrid.h:1278-1342 Accumulate (weights clamped to [0,1]). */
void vgo_accumulate(const vgo_grid *g, float *f, int ncomp, const float p[3],
                    const float *item) {
    uint32_t idx[3];
    indices_of_position(g, p, idx);
    float t[3], o[3];
    for (int a = 0; a < 3; ++a) {
        const float cmin = g->min[a] + (float)idx[a] * g->cell[a];
        t[a] = clamp01((p[a] - cmin) * g->cpe[a]);
        o[a] = 1.0f - t[a];
    }
    const size_t nx = g->np[0], nxy = (size_t)g->np[0] * g->np[1];
    const size_t o000 = idx[0] + nx * (idx[1] + (size_t)g->np[1] * idx[2]);
    const size_t offs[8] = {o000, o000 + 1, o000 + nx, o000 + nx + 1,
                            o000 + nxy, o000 + nxy + 1, o000 + nxy + nx,
                            o000 + nxy + nx + 1};
    const float w[8] = {o[0] * o[1] * o[2], t[0] * o[1] * o[2], o[0] * t[1] * o[2],
                        t[0] * t[1] * o[2], o[0] * o[1] * t[2], t[0] * o[1] * t[2],
                        o[0] * t[1] * t[2], t[0] * t[1] * t[2]};
    for (int k = 0; k < 8; ++k)
        for (int c = 0; c < ncomp; ++c) f[ncomp * offs[k] + c] += item[c] * w[k];
}

/* vortonSim.cpp:2372-2436 */
void vgo_splat_vorticity(const vgo_grid *g, const float *pos, const float *om,
                         size_t n, float vorton_radius, float scale, float *out) {
    memset(out, 0, vgo_grid_capacity(g) * 3 * sizeof(float));
    if (n == 0) return;
    const float cell_volume = g->cell[0] * g->cell[1] * g->cell[2];
    const float four_pi_over3 = (4.0f * kPi) / 3.0f;
    const float vorton_volume = four_pi_over3 * (vorton_radius * vorton_radius * vorton_radius);
    const float vc = scale * vorton_volume / cell_volume;
    for (size_t i = 0; i < n; ++i) {
        const float item[3] = {(om[3 * i] * 2.0f) * vc, (om[3 * i + 1] * 2.0f) * vc,
                               (om[3 * i + 2] * 2.0f) * vc};
        vgo_accumulate(g, out, 3, pos + 3 * i, item);
    }
}

/* pclOpPopulateVelocityGrid.cpp:33-87 */
void vgo_populate_velocity_grid(const vgo_grid *g, const float *pos,
                                const float *vel, size_t n, float *out) {
    const size_t cap = vgo_grid_capacity(g);
    memset(out, 0, cap * 3 * sizeof(float));
    float *wsum = (float *)calloc(cap, sizeof(float));
    const float one = 1.0f;
    for (size_t i = 0; i < n; ++i) {
        vgo_accumulate(g, out, 3, pos + 3 * i, vel + 3 * i);
        vgo_accumulate(g, wsum, 1, pos + 3 * i, &one);
    }
    for (size_t o = 0; o < cap; ++o) {
        if (wsum[o] > 0.0f) {
            const float inv = 1.0f / wsum[o];
            for (int c = 0; c < 3; ++c) out[3 * o + c] *= inv;
        }
    }
    free(wsum);
}

/* uniformGrid.h:1453-1585 DownSampleSlice */
void vgo_downsample(const vgo_grid *lo, float *lf, const vgo_grid *hi,
                    const float *hf, int accurate) {
    static const float mult[4] = {8.0f, 4.0f, 2.0f, 1.0f};
    uint32_t dims[3];
    for (int a = 0; a < 3; ++a) dims[a] = (hi->np[a] - 1) / (lo->np[a] - 1);
    const size_t hnx = hi->np[0], hnxy = (size_t)hi->np[0] * hi->np[1];
    for (uint32_t iz = 0; iz < lo->np[2]; ++iz)
        for (uint32_t iy = 0; iy < lo->np[1]; ++iy)
            for (uint32_t ix = 0; ix < lo->np[0]; ++ix) {
                const size_t lo_off = ix + (size_t)lo->np[0] * (iy + (size_t)lo->np[1] * iz);
                float *r = lf + 3 * lo_off;
                const uint32_t cm[3] = {ix * dims[0], iy * dims[1], iz * dims[2]};
                if (!accurate) {
                    const size_t ho = cm[0] + hnx * cm[1] + hnxy * cm[2];
                    for (int c = 0; c < 3; ++c) r[c] = hf[3 * ho + c];
                    continue;
                }
                float acc[3] = {0.f, 0.f, 0.f};
                const uint32_t sh[3] = {dims[0] / 2, dims[1] / 2, dims[2] / 2};
                for (uint32_t i2 = 0; i2 < dims[2] + sh[2]; ++i2) {
                    const int s2 = (int)i2 - (int)sh[2];
                    const uint32_t h2 = cm[2] + (uint32_t)s2;
                    if (h2 >= hi->np[2]) continue;
                    for (uint32_t i1 = 0; i1 < dims[1] + sh[1]; ++i1) {
                        const int s1 = (int)i1 - (int)sh[1];
                        const uint32_t h1 = cm[1] + (uint32_t)s1;
                        if (h1 >= hi->np[1]) continue;
                        for (uint32_t i0 = 0; i0 < dims[0] + sh[0]; ++i0) {
                            const int s0 = (int)i0 - (int)sh[0];
                            const uint32_t h0 = cm[0] + (uint32_t)s0;
                            if (h0 >= hi->np[0]) continue;
                            const unsigned md = (unsigned)(abs(s0) + abs(s1) + abs(s2));
                            const size_t ho = h0 + hnx * h1 + hnxy * h2;
                            for (int c = 0; c < 3; ++c) acc[c] += hf[3 * ho + c] * mult[md];
                        }
                    }
                }
                const float inv = 1.0f / 64.0f;
                for (int c = 0; c < 3; ++c) r[c] = acc[c] * inv;
            }
}

/* uniformGrid.h:1637-1700 UpSampleSlice, serial semantics (one slice
 * covering all z). */
void vgo_upsample(const vgo_grid *hi, float *hf, const vgo_grid *lo,
                  const float *lf, int interior_only) {
    float pmin[3], sp[3];
    for (int a = 0; a < 3; ++a) {
        const float adj = (fabsf(hi->min[a]) * 2.0f) * FLT_EPSILON;
        pmin[a] = hi->min[a] + adj;
        sp[a] = hi->cell[a] * (1.0f - 4.0f * FLT_EPSILON);
    }
    uint32_t lo_i[3] = {0, 0, 0}, hi_i[3] = {hi->np[0], hi->np[1], hi->np[2]};
    if (interior_only) {
        for (int a = 0; a < 3; ++a) {
            lo_i[a] = 1;
            hi_i[a] = (hi->np[a] > 1 ? hi->np[a] : 1) - 1;
        }
    }
    for (uint32_t iz = lo_i[2]; iz < hi_i[2]; ++iz) {
        float p[3];
        p[2] = pmin[2] + (float)iz * sp[2];
        for (uint32_t iy = lo_i[1]; iy < hi_i[1]; ++iy) {
            p[1] = pmin[1] + (float)iy * sp[1];
            for (uint32_t ix = lo_i[0]; ix < hi_i[0]; ++ix) {
                p[0] = pmin[0] + (float)ix * sp[0];
                const size_t off = ix + (size_t)hi->np[0] * (iy + (size_t)hi->np[1] * iz);
                vgo_interpolate(lo, lf, p, hf + 3 * off);
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* uniformGridMath.cpp:1116-1188 StepTowardVectorPoissonSolution (Dirichlet):
 * rows whose (iy+iz) parity matches the colour, x sequential in place. */
void vgo_sor_pass(const vgo_grid *g, float *s, const float *lap, int color) {
    const float rsx = 1.0f / g->cell[0], rsy = 1.0f / g->cell[1];
    const float rsz = g->cell[2] > FLT_EPSILON ? 1.0f / g->cell[2] : 0.0f;
    const float r2x = rsx * rsx, r2y = rsy * rsy, r2z = rsz * rsz;
    const float h = 0.5f / (r2x + r2y + r2z);
    const float relax = 1.72f, omr = 1.0f - relax;
    const size_t nx = g->np[0], ny = g->np[1], nz = g->np[2], nxy = nx * ny;
    for (size_t iz = 1; iz + 1 < nz; ++iz) {
        const size_t yshift = (iz % 2 == (size_t)color) ? 1 : 0;
        for (size_t iy = 1 + yshift; iy + 1 < ny; iy += 2) {
            const size_t row = iy * nx + iz * nxy;
            for (size_t ix = 1; ix + 1 < nx; ++ix) {
                const size_t o = ix + row;
                for (int c = 0; c < 3; ++c) {
                    const float v = ((s[3 * (o + 1) + c] + s[3 * (o - 1) + c]) * r2x +
                                     (s[3 * (o + nx) + c] + s[3 * (o - nx) + c]) * r2y +
                                     (s[3 * (o + nxy) + c] + s[3 * (o - nxy) + c]) * r2z -
                                     lap[3 * o + c]) * h;
                    s[3 * o + c] = omr * s[3 * o + c] + relax * v;
                }
            }
        }
    }
}

/* uniformGridMath.cpp:1369-1405 */
void vgo_solve_vector_poisson(const vgo_grid *g, float *soln, const float *lap,
                              size_t num_steps) {
    size_t maxdim = g->np[0];
    if (g->np[1] > maxdim) maxdim = g->np[1];
    if (g->np[2] > maxdim) maxdim = g->np[2];
    const size_t iters = num_steps > 0 ? num_steps : 2 * maxdim;
    for (size_t it = 0; it < iters; ++it) {
        vgo_sor_pass(g, soln, lap, 0);
        vgo_sor_pass(g, soln, lap, 1);
    }
}

/* vortonSim.cpp:2523-2571 (VORTON_SIM_USE_MULTI_GRID branch). */
void vgo_mg_cascade(const vgo_grid *g, float *a0, const float *negw0,
                    size_t steps) {
    vgo_nested ng;
    vgo_nested_init(&ng, g);
    float *A[VGO_MAX_LAYERS], *W[VGO_MAX_LAYERS];
    A[0] = a0;
    W[0] = (float *)negw0;
    for (uint32_t l = 1; l < ng.depth; ++l) {
        A[l] = (float *)calloc(vgo_grid_capacity(&ng.layer[l]) * 3, sizeof(float));
        W[l] = (float *)calloc(vgo_grid_capacity(&ng.layer[l]) * 3, sizeof(float));
    }
    vgo_solve_vector_poisson(&ng.layer[0], A[0], W[0], steps);
    uint32_t max_valid = 0;
    for (uint32_t l = 1; l < ng.depth; ++l) {
        const uint32_t *np = ng.layer[l].np;
        uint32_t md = np[0];
        if (np[1] < md) md = np[1];
        if (np[2] < md) md = np[2];
        if (md > 2) {
            vgo_downsample(&ng.layer[l], W[l], &ng.layer[l - 1], W[l - 1], 0);
            vgo_downsample(&ng.layer[l], A[l], &ng.layer[l - 1], A[l - 1], 0);
            vgo_solve_vector_poisson(&ng.layer[l], A[l], W[l], steps);
        } else {
            max_valid = l - 1;
            break;
        }
    }
    for (uint32_t l = max_valid; l >= 1; --l) {
        vgo_upsample(&ng.layer[l - 1], A[l - 1], &ng.layer[l], A[l], 1);
        vgo_solve_vector_poisson(&ng.layer[l - 1], A[l - 1], W[l - 1], steps);
    }
    for (uint32_t l = 1; l < ng.depth; ++l) { free(A[l]); free(W[l]); }
}

/* uniformGridMath.cpp:304-451 ComputeJacobian fused with :207-232
 * ComputeCurlFromJacobian. */
void vgo_curl(const vgo_grid *g, const float *v, float *curl) {
    const float rx = 1.0f / g->cell[0], ry = 1.0f / g->cell[1];
    const float rz = g->cell[2] > FLT_EPSILON ? 1.0f / g->cell[2] : 0.0f;
    const float r[3] = {rx, ry, rz};
    const float hr[3] = {0.5f * rx, 0.5f * ry, 0.5f * rz};
    const size_t n[3] = {g->np[0], g->np[1], g->np[2]};
    const size_t stride[3] = {1, n[0], n[0] * n[1]};
    for (size_t iz = 0; iz < n[2]; ++iz)
        for (size_t iy = 0; iy < n[1]; ++iy)
            for (size_t ix = 0; ix < n[0]; ++ix) {
                const size_t idx[3] = {ix, iy, iz};
                const size_t o = ix + n[0] * iy + n[0] * n[1] * iz;
                float j[3][3]; /* j[a][b] = d v_b / d a */
                for (int a = 0; a < 3; ++a) {
                    const size_t s = stride[a];
                    for (int b = 0; b < 3; ++b) {
                        float d;
                        if (idx[a] == 0)
                            d = (v[3 * (o + s) + b] - v[3 * o + b]) * r[a];
                        else if (idx[a] == n[a] - 1)
                            d = (v[3 * o + b] - v[3 * (o - s) + b]) * r[a];
                        else
                            d = (v[3 * (o + s) + b] - v[3 * (o - s) + b]) * hr[a];
                        j[a][b] = d;
                    }
                }
                curl[3 * o] = j[1][2] - j[2][1];
                curl[3 * o + 1] = j[2][0] - j[0][2];
                curl[3 * o + 2] = j[0][1] - j[1][0];
            }
}

/* uniformGridMath.cpp:304-451.  The reference computes the interior with
 * central differences, then overwrites the six faces (in the order -x, -y,
 * -z, +x, +y, +z) with the per-point rule: one-sided 1/h at an index 0 or
 * n-1 along that axis, central 1/(2h) otherwise.  Every face pass applies the
 * same per-point rule, so the result is that rule at every point. */
void vgo_jacobian(const vgo_grid *g, const float *v, float *jac) {
    const float rx = 1.0f / g->cell[0], ry = 1.0f / g->cell[1];
    const float rz = g->cell[2] > FLT_EPSILON ? 1.0f / g->cell[2] : 0.0f;
    const float r[3] = {rx, ry, rz};
    const float hr[3] = {0.5f * rx, 0.5f * ry, 0.5f * rz};
    const size_t n[3] = {g->np[0], g->np[1], g->np[2]};
    const size_t stride[3] = {1, n[0], n[0] * n[1]};
    for (size_t iz = 0; iz < n[2]; ++iz)
        for (size_t iy = 0; iy < n[1]; ++iy)
            for (size_t ix = 0; ix < n[0]; ++ix) {
                const size_t idx[3] = {ix, iy, iz};
                const size_t o = ix + n[0] * iy + n[0] * n[1] * iz;
                for (int a = 0; a < 3; ++a) {
                    const size_t s = stride[a];
                    size_t lo = o - s, hi = o + s;
                    float scale = hr[a];
                    if (idx[a] == 0) { lo = o; scale = r[a]; }
                    else if (idx[a] == n[a] - 1) { hi = o; scale = r[a]; }
                    for (int b = 0; b < 3; ++b)
                        jac[9 * o + 3 * a + b] = (v[3 * hi + b] - v[3 * lo + b]) * scale;
                }
            }
}

/* Trilinear interpolation of an ncomp-component grid, the nested-lerp
 * order of uniformGrid.h:1137-1144 applied to every component. */
static void interpolate_n(const vgo_grid *g, const float *f, int ncomp, const float p[3],
                          float *out) {
    uint32_t idx[3];
    indices_of_position(g, p, idx);
    float t[3], o[3];
    for (int a = 0; a < 3; ++a) {
        const float cmin = g->min[a] + (float)idx[a] * g->cell[a];
        t[a] = (p[a] - cmin) * g->cpe[a];
        o[a] = 1.0f - t[a];
    }
    const size_t nx = g->np[0], nxy = (size_t)g->np[0] * g->np[1];
    const size_t o000 = idx[0] + nx * (idx[1] + (size_t)g->np[1] * idx[2]);
    const size_t c[8] = {o000, o000 + 1, o000 + nx, o000 + nx + 1,
                         o000 + nxy, o000 + nxy + 1, o000 + nxy + nx, o000 + nxy + nx + 1};
    for (int k = 0; k < ncomp; ++k) {
#define F(i) f[(size_t)ncomp * c[i] + k]
        out[k] = ((o[0] * F(0) + t[0] * F(1)) * o[1] + (o[0] * F(2) + t[0] * F(3)) * t[1]) * o[2] +
                 ((o[0] * F(4) + t[0] * F(5)) * o[1] + (o[0] * F(6) + t[0] * F(7)) * t[1]) * t[2];
#undef F
    }
}

/* vortonSim.cpp:2748-2820 (transpose formulation, :2800-2801). */
void vgo_stretch_tilt(const vgo_grid *g, const float *jac, const float *pos,
                      size_t pos_stride, float *angvel, size_t ang_stride,
                      size_t n, float dt) {
    if (g->extent[0] == 0.0f || g->extent[1] == 0.0f || g->extent[2] == 0.0f) return;
    for (size_t i = 0; i < n; ++i) {
        float j[9];
        interpolate_n(g, jac, 9, pos + i * pos_stride, j);
        float *w = angvel + i * ang_stride;
        float st[3];
        for (int a = 0; a < 3; ++a)  /* Vec3 dot, scalar build: x*x + y*y + z*z */
            st[a] = j[3 * a] * w[0] + j[3 * a + 1] * w[1] + j[3 * a + 2] * w[2];
        for (int a = 0; a < 3; ++a) w[a] += st[a] * 0.5f * dt;
    }
}

/* Cody & Waite (1980) range reduction with the 16-bit split of ln 2. */
float vgo_pse_exp(float x) {
    if (x != x) return x;
    if (x < -87.0f) return 0.0f;
    const float k = rintf(x * 1.44269504f);
    const float r = (x - k * 0.693145751953125f) - k * 1.42860677e-06f;
    const float p =
        1.0f + r * (1.0f + r * (0.5f + r * (0.166666672f + r * (0.0416666679f +
        r * (0.00833333377f + r * (0.00138888892f + r * 0.000198412701f))))));
    const int32_t bits = ((int32_t)k + 127) << 23;
    float scale;
    memcpy(&scale, &bits, 4);
    return p * scale;
}

/* vortonSim.cpp:3618-3668 ExchangeVorticityOrMergeVortons (merging off). */
static void pse_exchange(const float *ph, float r2, float *wh, const float *pt, float *wt,
                         float nu_dt2) {
    const float sx = ph[0] - pt[0], sy = ph[1] - pt[1], sz = ph[2] - pt[2];
    const float d2 = sx * sx + sy * sy + sz * sz;
    const float law = vgo_pse_exp(-d2 / r2);
    float f = nu_dt2 * law;
    f = f < -0.5f ? -0.5f : (f > 0.5f ? 0.5f : f);
    for (int c = 0; c < 3; ++c) {
        const float e = (wh[c] - wt[c]) * f;
        wh[c] -= e;
        wt[c] += e;
    }
}

void vgo_pse_diffuse(const vgo_grid *g, const float *pos, float *angvel,
                     const float *size, float uniform_size,
                     const unsigned char *alive, size_t n, float viscosity,
                     float dt) {
    const size_t nx = g->np[0], ny = g->np[1], nz = g->np[2], nxy = nx * ny;
    const size_t cap = nxy * nz;
    if (n == 0 || nx < 2 || ny < 2 || nz < 2) return;
    /* PartitionParticles: per cell, vorton indices in increasing order */
    uint32_t *count = (uint32_t *)calloc(cap, sizeof(uint32_t));
    uint32_t *start = (uint32_t *)calloc(cap + 1, sizeof(uint32_t));
    uint32_t *list = (uint32_t *)malloc(n * sizeof(uint32_t));
    size_t *cell = (size_t *)malloc(n * sizeof(size_t));
    for (size_t v = 0; v < n; ++v) {
        uint32_t idx[3];
        indices_of_position(g, pos + 3 * v, idx);
        cell[v] = idx[0] + nx * (idx[1] + ny * (size_t)idx[2]);
        ++count[cell[v]];
    }
    for (size_t c = 0; c < cap; ++c) start[c + 1] = start[c] + count[c];
    memset(count, 0, cap * sizeof(uint32_t));
    for (size_t v = 0; v < n; ++v) list[start[cell[v]] + count[cell[v]]++] = (uint32_t)v;
    const float nu_dt2 = 2.0f * viscosity * dt;  /* 2.0f * mViscosity * timeStep */
    const float dissipation = viscosity * dt;
    const size_t fwd[7] = {1, nx, 1 + nx, nxy, 1 + nxy, nx + nxy, 1 + nx + nxy};
    for (size_t phase = 0; phase < 2; ++phase)
        for (size_t z = phase; z + 1 < nz; z += 2)
            for (size_t y = 0; y + 1 < ny; ++y)
                for (size_t x = 0; x + 1 < nx; ++x) {
                    const size_t c = x + nx * y + nxy * z;
                    for (uint32_t i = start[c]; i < start[c + 1]; ++i) {
                        const uint32_t h = list[i];
                        if (alive && !alive[h]) continue;
                        const float sh = size ? size[h] : uniform_size;
                        const float range = 4.0f * sh;
                        const float r2 = range * range;
                        float *wh = angvel + 3 * (size_t)h;
                        for (uint32_t j = i + 1; j < start[c + 1]; ++j)
                            pse_exchange(pos + 3 * (size_t)h, r2, wh, pos + 3 * (size_t)list[j],
                                         angvel + 3 * (size_t)list[j], nu_dt2);
                        for (int k = 0; k < 7; ++k) {
                            const size_t cn = c + fwd[k];
                            for (uint32_t j = start[cn]; j < start[cn + 1]; ++j)
                                pse_exchange(pos + 3 * (size_t)h, r2, wh,
                                             pos + 3 * (size_t)list[j],
                                             angvel + 3 * (size_t)list[j], nu_dt2);
                        }
                        for (int a = 0; a < 3; ++a) wh[a] -= wh[a] * dissipation;
                    }
                }
    free(count); free(start); free(list); free(cell);
}

/* e = f (Omega_h - Omega_t) with f from h's size, as in pse_exchange. */
static void pse_pair(const float *ph, float r2h, const float *wh, const float *pt,
                     const float *wt, float nu_dt2, float e[3]) {
    const float sx = ph[0] - pt[0], sy = ph[1] - pt[1], sz = ph[2] - pt[2];
    const float d2 = sx * sx + sy * sy + sz * sz;
    float f = nu_dt2 * vgo_pse_exp(-d2 / r2h);
    f = f < -0.5f ? -0.5f : (f > 0.5f ? 0.5f : f);
    for (int c = 0; c < 3; ++c) e[c] = (wh[c] - wt[c]) * f;
}

void vgo_pse_diffuse_simultaneous(const vgo_grid *g, const float *pos, float *angvel,
                                  const float *size, float uniform_size,
                                  const unsigned char *alive, size_t n,
                                  float viscosity, float dt) {
    const size_t nx = g->np[0], ny = g->np[1], nz = g->np[2], nxy = nx * ny;
    const size_t cap = nxy * nz;
    if (n == 0 || nx < 2 || ny < 2 || nz < 2) return;
    uint32_t *count = (uint32_t *)calloc(cap, sizeof(uint32_t));
    uint32_t *start = (uint32_t *)calloc(cap + 1, sizeof(uint32_t));
    uint32_t *list = (uint32_t *)malloc(n * sizeof(uint32_t));
    uint32_t *slot = (uint32_t *)malloc(n * sizeof(uint32_t));
    size_t *cell = (size_t *)malloc(n * sizeof(size_t));
    uint32_t (*cidx)[3] = malloc(n * sizeof *cidx);
    float *w0 = (float *)malloc(3 * n * sizeof(float));
    memcpy(w0, angvel, 3 * n * sizeof(float));
    for (size_t v = 0; v < n; ++v) {
        indices_of_position(g, pos + 3 * v, cidx[v]);
        cell[v] = cidx[v][0] + nx * (cidx[v][1] + ny * (size_t)cidx[v][2]);
        ++count[cell[v]];
    }
    for (size_t c = 0; c < cap; ++c) start[c + 1] = start[c] + count[c];
    memset(count, 0, cap * sizeof(uint32_t));
    for (size_t v = 0; v < n; ++v) {
        slot[v] = start[cell[v]] + count[cell[v]]++;
        list[slot[v]] = (uint32_t)v;
    }
    const float nu_dt2 = 2.0f * viscosity * dt;
    const float dissipation = viscosity * dt;
    const int d[7][3] = {{1, 0, 0}, {0, 1, 0}, {1, 1, 0}, {0, 0, 1}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
    for (size_t i = 0; i < n; ++i) {
        const size_t c = cell[i];
        const float *pi = pos + 3 * i, *wi = w0 + 3 * i;
        float acc[3] = {0.f, 0.f, 0.f}, e[3];
        for (int k = 0; k < 7; ++k) {  /* received from backward cells */
            if ((int)cidx[i][0] < d[k][0] || (int)cidx[i][1] < d[k][1] ||
                (int)cidx[i][2] < d[k][2])
                continue;
            const size_t cb = c - (d[k][0] + nx * d[k][1] + nxy * d[k][2]);
            for (uint32_t j = start[cb]; j < start[cb + 1]; ++j) {
                const uint32_t h = list[j];
                if (alive && !alive[h]) continue;
                const float sh = 4.0f * (size ? size[h] : uniform_size);
                pse_pair(pos + 3 * (size_t)h, sh * sh, w0 + 3 * (size_t)h, pi, wi, nu_dt2, e);
                for (int a = 0; a < 3; ++a) acc[a] += e[a];
            }
        }
        for (uint32_t j = start[c]; j < slot[i]; ++j) {  /* received from earlier in the cell */
            const uint32_t h = list[j];
            if (alive && !alive[h]) continue;
            const float sh = 4.0f * (size ? size[h] : uniform_size);
            pse_pair(pos + 3 * (size_t)h, sh * sh, w0 + 3 * (size_t)h, pi, wi, nu_dt2, e);
            for (int a = 0; a < 3; ++a) acc[a] += e[a];
        }
        const int live = !alive || alive[i];
        if (live) {
            const float si = 4.0f * (size ? size[i] : uniform_size);
            const float r2 = si * si;
            for (uint32_t j = slot[i] + 1; j < start[c + 1]; ++j) {
                const uint32_t t = list[j];
                pse_pair(pi, r2, wi, pos + 3 * (size_t)t, w0 + 3 * (size_t)t, nu_dt2, e);
                for (int a = 0; a < 3; ++a) acc[a] -= e[a];
            }
            for (int k = 0; k < 7; ++k) {
                const size_t cf = c + (d[k][0] + nx * d[k][1] + nxy * d[k][2]);
                for (uint32_t j = start[cf]; j < start[cf + 1]; ++j) {
                    const uint32_t t = list[j];
                    pse_pair(pi, r2, wi, pos + 3 * (size_t)t, w0 + 3 * (size_t)t, nu_dt2, e);
                    for (int a = 0; a < 3; ++a) acc[a] -= e[a];
                }
            }
        }
        float *wo = angvel + 3 * i;
        for (int a = 0; a < 3; ++a) {
            wo[a] = wi[a] + acc[a];
            if (live) wo[a] -= wo[a] * dissipation;
        }
    }
    free(count); free(start); free(list); free(slot); free(cell); free(cidx); free(w0);
}

void vgo_brick_of_bytes(const vgo_grid *g, const float *field, unsigned char *planes,
                        float stats[7]) {
    const size_t cap = (size_t)g->np[0] * g->np[1] * g->np[2];
    float mn[3] = {field[0], field[1], field[2]}, mx[3] = {field[0], field[1], field[2]};
    for (size_t o = 0; o < cap; ++o) {
        const float *r = field + 3 * o;
        /* Min2(a, b) = a < b ? a : b with Vec3 '<' true only when all three
         * components are strictly less; likewise Max2 with '>' */
        if (!(mn[0] < r[0] && mn[1] < r[1] && mn[2] < r[2]))
            for (int a = 0; a < 3; ++a) mn[a] = r[a];
        if (!(mx[0] > r[0] && mx[1] > r[1] && mx[2] > r[2]))
            for (int a = 0; a < 3; ++a) mx[a] = r[a];
    }
    float ext[3], range[3];
    for (int a = 0; a < 3; ++a) {
        ext[a] = -mn[a] > mx[a] ? -mn[a] : mx[a];
        const float d = mx[a] - mn[a];
        range[a] = FLT_MIN > d ? FLT_MIN : d;
    }
    const float magmax = sqrtf(ext[0] * ext[0] + ext[1] * ext[1] + ext[2] * ext[2]);
    const float almost256 = 256.0f * (1.0f - FLT_EPSILON);
    for (size_t o = 0; o < cap; ++o) {
        const float *v = field + 3 * o;
        const float mag = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        for (int a = 0; a < 3; ++a) {
            const float t = (v[a] - mn[a]) / range[a];
            planes[a * cap + o] = (unsigned char)(int)(t * almost256);
        }
        planes[3 * cap + o] = (unsigned char)(int)(almost256 * mag / (magmax + FLT_MIN));
    }
    for (int a = 0; a < 3; ++a) {
        stats[a] = mn[a];
        stats[3 + a] = mx[a];
    }
    stats[6] = magmax;
}

/* FP64 true relative residual of the 7-point Dirichlet Laplacian over the
 * interior: ||f - L u|| / ||f||, u = hi (+ lo when lo is not NULL, the
 * compensated pair of a solver that carries its solution in two floats),
 * each value summed in double. */
double vgo_residual_rel_pair(const vgo_grid *g, const float *hi, const float *lo,
                             const float *f) {
    const double rx = 1.0 / g->cell[0], ry = 1.0 / g->cell[1], rz = 1.0 / g->cell[2];
    const double r2[3] = {rx * rx, ry * ry, rz * rz};
    const size_t nx = g->np[0], ny = g->np[1], nz = g->np[2], nxy = nx * ny;
#define UP(o) ((double)hi[3 * (o) + c] + (lo ? (double)lo[3 * (o) + c] : 0.0))
    double num = 0, den = 0;
    for (size_t iz = 1; iz + 1 < nz; ++iz)
        for (size_t iy = 1; iy + 1 < ny; ++iy)
            for (size_t ix = 1; ix + 1 < nx; ++ix) {
                const size_t o = ix + iy * nx + iz * nxy;
                for (int c = 0; c < 3; ++c) {
                    const double uc = UP(o);
                    const double lap = (UP(o + 1) + UP(o - 1) - 2 * uc) * r2[0] +
                                       (UP(o + nx) + UP(o - nx) - 2 * uc) * r2[1] +
                                       (UP(o + nxy) + UP(o - nxy) - 2 * uc) * r2[2];
                    const double res = (double)f[3 * o + c] - lap;
                    num += res * res;
                    den += (double)f[3 * o + c] * f[3 * o + c];
                }
            }
#undef UP
    return den > 0 ? sqrt(num / den) : sqrt(num);
}

double vgo_residual_rel(const vgo_grid *g, const float *u, const float *f) {
    const double rx = 1.0 / g->cell[0], ry = 1.0 / g->cell[1], rz = 1.0 / g->cell[2];
    const double r2[3] = {rx * rx, ry * ry, rz * rz};
    const size_t nx = g->np[0], ny = g->np[1], nz = g->np[2], nxy = nx * ny;
    double num = 0, den = 0;
    for (size_t iz = 1; iz + 1 < nz; ++iz)
        for (size_t iy = 1; iy + 1 < ny; ++iy)
            for (size_t ix = 1; ix + 1 < nx; ++ix) {
                const size_t o = ix + iy * nx + iz * nxy;
                for (int c = 0; c < 3; ++c) {
                    const double uc = u[3 * o + c];
                    const double lap =
                        ((double)u[3 * (o + 1) + c] + u[3 * (o - 1) + c] - 2 * uc) * r2[0] +
                        ((double)u[3 * (o + nx) + c] + u[3 * (o - nx) + c] - 2 * uc) * r2[1] +
                        ((double)u[3 * (o + nxy) + c] + u[3 * (o - nxy) + c] - 2 * uc) * r2[2];
                    const double res = (double)f[3 * o + c] - lap;
                    num += res * res;
                    den += (double)f[3 * o + c] * f[3 * o + c];
                }
            }
    return den > 0 ? sqrt(num / den) : sqrt(num);
}

/* ------------------------------------------------------------------------ */
/* pclOpAdvect.cpp:42-81 */
void vgo_assign_velocity_from_field(const vgo_grid *g, const float *field,
                                    float *pos, float *vel, size_t stride,
                                    size_t n, float gain) {
    (void)pos;
    if (g->extent[0] == 0.0f && g->extent[1] == 0.0f && g->extent[2] == 0.0f) return;
    for (size_t i = 0; i < n; ++i) {
        float v[3];
        vgo_interpolate(g, field, pos + i * stride, v);
        float *pv = vel + i * stride;
        if (1.0f == gain) {
            pv[0] = v[0]; pv[1] = v[1]; pv[2] = v[2];
        } else {
            for (int c = 0; c < 3; ++c) {
                const float d = v[c] - pv[c];
                pv[c] += gain * d;
            }
        }
    }
}

/* pclOpEvolve.cpp:36-55 */
void vgo_evolve(float *pos, const float *vel, size_t stride, size_t n, float dt) {
    for (size_t i = 0; i < n; ++i) {
        float *p = pos + i * stride;
        const float *v = vel + i * stride;
        for (int c = 0; c < 3; ++c) p[c] += v[c] * dt;
    }
}
