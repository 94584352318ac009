/*
 * CPU oracle for the Gravity particle step: TEST INFRASTRUCTURE ONLY.
 * See oracle.h for the layout, the canonical orders and who may use it.
 *
 * Built with -ffp-contract=off so every float expression rounds as written:
 * the shaders' `precise` contact predicate and the HIP path (also built
 * without contraction) see the same operations.
 *
 * Citations: S/ = src/main/resources/shaders/, K/ = src/main/kotlin/dev/gravity/.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MAXIMUM_UNIT 0.99999994f

/* clamp as fmin(fmax(v, lo), hi): a NaN lands on lo, as on the GPU. */
static float clampf(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }
static int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
static float fractf(float v) { return v - floorf(v); }
/* GLSL mod(x, y) = x - y * floor(x / y). */
static float glsl_mod(float x, float y) { return x - y * floorf(x / y); }
/* Float -> int of an already floored value, clamped in float first so that
 * non-finite inputs land on a valid index on every platform. */
static int floor_index(float v, int size) {
    float f = floorf(v);
    if (!(f >= 0.0f)) return 0;
    if (f > (float)(size - 1)) return size - 1;
    return (int)f;
}

/* ------------------------------------------------------------------------ */
/* a: spatial sort (S/spatial_sort_key.glsl:7-25, K/GravityPipeline.kt:1647) */

uint32_t orc_spread_bits(uint32_t v) {
    v &= 0xffffu;
    v = (v | (v << 8)) & 0x00ff00ffu;
    v = (v | (v << 4)) & 0x0f0f0f0fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

uint32_t orc_sort_key(float x, float y, int wb, int hb) {
    int sx = 1 << wb, sy = 1 << hb;
    uint32_t tx = (uint32_t)floor_index(x * (float)sx, sx);
    uint32_t ty = (uint32_t)floor_index(y * (float)sy, sy);
    int common = wb < hb ? wb : hb;
    uint32_t low = (1u << common) - 1u;
    uint32_t morton = orc_spread_bits(tx & low) | (orc_spread_bits(ty & low) << 1);
    return morton | (((tx | ty) >> common) << (2 * common));
}

int orc_sort_axis_bits(int extent) {
    int tiles = extent / 4;
    if (tiles < 1) tiles = 1;
    int bits = 31 - __builtin_clz((unsigned)tiles);
    return bits > 10 ? 10 : bits;
}

void orc_spatial_sort(int n, int width, int height, float *bodies, float *accel,
                      float *materials, uint32_t *dest) {
    int wb = orc_sort_axis_bits(width), hb = orc_sort_axis_bits(height);
    uint32_t tiles = 1u << (wb + hb);
    uint32_t *count = calloc(tiles + 1, sizeof(uint32_t));
    uint32_t *key = malloc((size_t)n * sizeof(uint32_t));
    for (int i = 0; i < n; i++) {
        key[i] = orc_sort_key(bodies[4 * i], bodies[4 * i + 1], wb, hb);
        count[key[i] + 1]++;
    }
    for (uint32_t t = 0; t < tiles; t++) count[t + 1] += count[t];
    for (int i = 0; i < n; i++) dest[i] = count[key[i]]++;
    float *b = malloc((size_t)n * 16), *a = malloc((size_t)n * 8), *m = malloc((size_t)n * 16);
    for (int i = 0; i < n; i++) {
        memcpy(b + 4 * dest[i], bodies + 4 * i, 16);
        memcpy(a + 2 * dest[i], accel + 2 * i, 8);
        memcpy(m + 4 * dest[i], materials + 4 * i, 16);
    }
    memcpy(bodies, b, (size_t)n * 16);
    memcpy(accel, a, (size_t)n * 8);
    memcpy(materials, m, (size_t)n * 16);
    free(b); free(a); free(m); free(key); free(count);
}

/* ------------------------------------------------------------------------ */
/* c: collisions                                                            */

/* K/ToroidalSpace.kt:21-22 */
int orc_collision_grid_cells(int extent, float diameter) {
    int c = (int)floorf((float)extent / diameter);
    return c < 1 ? 1 : c;
}

/* S/collision_crowd.glsl:2-15: the lower index goes first, no contraction. */
float orc_pair_distance_squared(uint32_t id, const float *pos, uint32_t other,
                                const float *opos, float vw, float vh, int bounded) {
    int own_first = id < other;
    const float *a = own_first ? pos : opos;
    const float *b = own_first ? opos : pos;
    float dx = a[0] * vw - b[0] * vw;
    float dy = a[1] * vh - b[1] * vh;
    if (!bounded) {
        dx -= floorf(dx / vw + 0.5f) * vw;
        dy -= floorf(dy / vh + 0.5f) * vh;
    }
    return dx * dx + dy * dy;
}

static int hard_pair(const orc_params *p, const float *materials, int a, int b) {
    if (!p->liquid) return 1;
    return materials[4 * a + 2] != 0.0f || materials[4 * b + 2] != 0.0f;
}

/* Broad phase: any exhaustive grid will do; cells are at least
 * diameter + margin wide so a 3x3 stencil holds every contact. */
typedef struct { int gw, gh; int *start; int *items; } cell_grid;

static void grid_build(cell_grid *g, int n, const float *bodies, float vw, float vh, float reach) {
    g->gw = (int)floorf(vw / reach); if (g->gw < 1) g->gw = 1;
    g->gh = (int)floorf(vh / reach); if (g->gh < 1) g->gh = 1;
    int cells = g->gw * g->gh;
    g->start = calloc((size_t)cells + 1, sizeof(int));
    g->items = malloc((size_t)(n > 0 ? n : 1) * sizeof(int));
    int *cell = malloc((size_t)(n > 0 ? n : 1) * sizeof(int));
    for (int i = 0; i < n; i++) {
        int cx = floor_index(bodies[4 * i] * (float)g->gw, g->gw);
        int cy = floor_index(bodies[4 * i + 1] * (float)g->gh, g->gh);
        cell[i] = cx + cy * g->gw;
        g->start[cell[i] + 1]++;
    }
    for (int c = 0; c < cells; c++) g->start[c + 1] += g->start[c];
    int *fill = malloc((size_t)cells * sizeof(int));
    memcpy(fill, g->start, (size_t)cells * sizeof(int));
    for (int i = 0; i < n; i++) g->items[fill[cell[i]]++] = i;
    free(fill); free(cell);
}

static void grid_free(cell_grid *g) { free(g->start); free(g->items); }

static int cmp_u32(const void *a, const void *b) {
    uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return x < y ? -1 : (x > y);
}

/* S/collision_cache_count.comp:61-95,130-285 (non-RIGID): every hard pair
 * with d^2 < D^2 (strict), listed in ascending index. */
void orc_contacts(const orc_params *p, int n, const float *bodies, const float *materials,
                  uint32_t *counts, uint32_t *lists, int cap) {
    float vw = (float)p->width, vh = (float)p->height;
    float d2 = p->diameter * p->diameter;
    float margin = 64.0f * (nextafterf(fmaxf(vw, vh), INFINITY) - fmaxf(vw, vh));
    cell_grid g;
    grid_build(&g, n, bodies, vw, vh, p->diameter + margin);
    uint32_t *tmp = malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < n; i++) {
        int k = 0;
        int no_hard = p->liquid && materials[4 * i + 2] == 0.0f;
        int cx = floor_index(bodies[4 * i] * (float)g.gw, g.gw);
        int cy = floor_index(bodies[4 * i + 1] * (float)g.gh, g.gh);
        int rx = g.gw >= 3 ? 1 : 0, ry = g.gh >= 3 ? 1 : 0;
        int visited[9]; int nv = 0;
        for (int dy = -1; dy <= 1; dy++) {
            for (int dx = -1; dx <= 1; dx++) {
                int x = cx + dx, y = cy + dy;
                if (p->bounded) {
                    if (x < 0 || y < 0 || x >= g.gw || y >= g.gh) continue;
                } else {
                    x = (x % g.gw + g.gw) % g.gw;
                    y = (y % g.gh + g.gh) % g.gh;
                }
                int c = x + y * g.gw, dup = 0;
                for (int v = 0; v < nv; v++) dup |= visited[v] == c;
                if (dup) continue;
                visited[nv++] = c;
                for (int t = g.start[c]; t < g.start[c + 1]; t++) {
                    int j = g.items[t];
                    if (j == i) continue;
                    if (no_hard && materials[4 * j + 2] == 0.0f) continue;
                    float dd = orc_pair_distance_squared((uint32_t)i, bodies + 4 * i, (uint32_t)j,
                                                         bodies + 4 * j, vw, vh, p->bounded);
                    if (dd < d2) tmp[k++] = (uint32_t)j;
                }
            }
        }
        (void)rx; (void)ry;
        qsort(tmp, (size_t)k, sizeof(uint32_t), cmp_u32);
        counts[i] = (uint32_t)k;
        if (lists) for (int s = 0; s < k && s < cap; s++) lists[(size_t)i * cap + s] = tmp[s];
    }
    free(tmp);
    grid_free(&g);
}

/* S/collision_cache_drift.comp:79-91 (sleep off). */
void orc_drift(const orc_params *p, int n, float *bodies, const float *accel,
               float kick_dt, float drift_dt) {
    if (kick_dt == 0.0f && drift_dt == 0.0f) return;
    for (int i = 0; i < n; i++) {
        float *b = bodies + 4 * i;
        if (kick_dt != 0.0f) {
            float h = 0.5f * kick_dt;
            b[2] += accel[2 * i] * h;
            b[3] += accel[2 * i + 1] * h;
        }
        b[0] += b[2] * drift_dt;
        b[1] += b[3] * drift_dt;
        if (p->bounded) {
            b[0] = clampf(b[0], 0.0f, MAXIMUM_UNIT);
            b[1] = clampf(b[1], 0.0f, MAXIMUM_UNIT);
        } else {
            b[0] = fractf(b[0]);
            b[1] = fractf(b[1]);
        }
    }
}

/* S/collision_cache_solve.comp:105-112 */
static void fallback_normal(uint32_t first, uint32_t second, float *nx, float *ny) {
    uint32_t low = first < second ? first : second;
    uint32_t high = first < second ? second : first;
    uint32_t hash = low * 1664525u + high * 1013904223u;
    /* The FP32 angle takes one of 65536 values; its cos/sin are the FP64
     * values rounded once (the HIP path reads the same table). */
    float angle = (float)(hash & 65535u) * (float)(6.28318530718 / 65536.0);
    float c = (float)cos((double)angle), s = (float)sin((double)angle);
    if (first < second) { *nx = c; *ny = s; } else { *nx = -c; *ny = -s; }
}

/* S/collision_material.glsl:128-135 */
static void friction_correction(float rvx, float rvy, float nx, float ny, float rvn,
                                float normal_change, float sliding_share, float friction,
                                float *ox, float *oy) {
    float tx = rvx - rvn * nx, ty = rvy - rvn * ny;
    float ts = sqrtf(tx * tx + ty * ty);
    if (ts <= 0.0f) { *ox = 0.0f; *oy = 0.0f; return; }
    float a = sliding_share * ts, b = friction * normal_change;
    float change = a < b ? a : b;
    float k = -change / ts;
    *ox = tx * k;
    *oy = ty * k;
}

/* S/collision_cache_solve.comp:55-86 */
static void wall_contact(const orc_params *p, float distance, float nx, float ny,
                         float vx, float vy, float weight, float friction, float eff_dt,
                         float *pcx, float *pcy, float *vcx, float *vcy, float *pressure,
                         int write_pressure) {
    float penetration = 0.5f * p->diameter - distance;
    if (penetration <= 0.0f) return;
    *pcx += nx * (penetration * weight);
    *pcy += ny * (penetration * weight);
    float ns = vx * nx + vy * ny;
    float e1 = 1.0f + p->restitution;
    if (ns < 0.0f) {
        float k = e1 * ns * weight;
        *vcx -= nx * k;
        *vcy -= ny * k;
    }
    if (friction > 0.0f) {
        float fx, fy;
        float approach = -ns > 0.0f ? -ns : 0.0f;
        friction_correction(vx, vy, nx, ny, ns, weight * (e1 * approach + penetration / eff_dt),
                            weight, friction, &fx, &fy);
        *vcx += fx;
        *vcy += fy;
    }
    if (write_pressure) {
        float approach = -ns > 0.0f ? -ns : 0.0f;
        *pressure += weight * (penetration / eff_dt + approach);
    }
}

static void accumulate_contact(const orc_params *p, uint32_t id, uint32_t other, const float *in,
                               const float *materials, const uint32_t *counts, float ppx, float ppy,
                               float vpx, float vpy, uint32_t own_contacts, float eff_dt,
                               int other_sleeps, float *pcx, float *pcy, float *vcx, float *vcy,
                               float *pressure, int write_pressure) {
    float vw = (float)p->width, vh = (float)p->height;
    if (!hard_pair(p, materials, (int)id, (int)other)) return;
    const float *o = in + 4 * other;
    float dx = ppx - o[0] * vw;
    float dy = ppy - o[1] * vh;
    if (!p->bounded) {
        dx -= floorf(dx / vw + 0.5f) * vw;
        dy -= floorf(dy / vh + 0.5f) * vh;
    }
    float d2 = dx * dx + dy * dy;
    float reach = p->diameter * p->diameter;
    if (!(d2 < reach)) return;
    float eps = p->coincident_eps;
    float distance = sqrtf(d2 > eps ? d2 : eps);
    float nx, ny;
    if (d2 > eps) { nx = dx / distance; ny = dy / distance; }
    else fallback_normal(id, other, &nx, &ny);
    float penetration = p->diameter - distance;
    uint32_t oc = counts[other];
    uint32_t crowd = own_contacts > oc ? own_contacts : oc;
    float pair_weight = 1.0f / (float)(crowd > 1u ? crowd : 1u);
    float m_own = materials[4 * id], m_other = materials[4 * other];
    /* A sleeping body stays where it is: this side takes the whole correction
     * (S/collision_cache_solve.comp:141-142,201). */
    float share = other_sleeps ? 1.0f : m_other / (m_own + m_other);
    float ovx = o[2] * vw, ovy = o[3] * vh;
    float rvx = vpx - ovx, rvy = vpy - ovy;
    float rvn = rvx * nx + rvy * ny;
    float push = share * penetration * pair_weight;
    *pcx += nx * push;
    *pcy += ny * push;
    float e1 = 1.0f + p->restitution;
    if (rvn < 0.0f) {
        float k = share * e1 * rvn * pair_weight;
        *vcx -= nx * k;
        *vcy -= ny * k;
    }
    float friction = sqrtf(materials[4 * id + 1] * materials[4 * other + 1]);
    if (friction > 0.0f) {
        float approach = -rvn > 0.0f ? -rvn : 0.0f;
        float fx, fy;
        friction_correction(rvx, rvy, nx, ny, rvn,
                            share * pair_weight * (e1 * approach + penetration / eff_dt),
                            share * pair_weight, friction, &fx, &fy);
        *vcx += fx;
        *vcy += fy;
    }
    if (write_pressure) {
        float approach = -rvn > 0.0f ? -rvn : 0.0f;
        *pressure += 0.5f * pair_weight * (penetration / eff_dt + approach);
    }
}

/* Sleeping tiles of a simulation (S/sleep_layout.glsl, K/SleepTiles.kt:37-107). */
typedef struct {
    int sleeping;             /* uSleep: enabled and past the wake steps */
    int enabled;              /* the final pass marks tiles */
    uint32_t stamp;
    int tiles_x, tiles_y;
    uint32_t *body_awake, *tile_awake, *tile_active;
    float disp2;              /* (sleepSpeedPixels * dt)^2 */
} sleep_ctx;

#define SLEEP_TILE_PIXELS 32
#define SLEEP_WAKE_STEPS 2

/* S/sleep_layout.glsl sleepTile */
static int sleep_tile(const sleep_ctx *sl, float x, float y) {
    float fx = floorf(x * (float)sl->tiles_x), fy = floorf(y * (float)sl->tiles_y);
    int tx = fx >= 0.0f ? (fx < (float)(sl->tiles_x - 1) ? (int)fx : sl->tiles_x - 1) : 0;
    int ty = fy >= 0.0f ? (fy < (float)(sl->tiles_y - 1) ? (int)fy : sl->tiles_y - 1) : 0;
    return tx + ty * sl->tiles_x;
}

static int sleep_awake(const sleep_ctx *sl, int i) {
    return !sl || !sl->sleeping || sl->body_awake[i] == sl->stamp;
}

/* S/collision_cache_solve.comp:274-478 (non-RIGID); sl (nullable) adds the
 * sleeping tiles: sleeping bodies are carried over, a sleeping partner takes
 * no share, and the final pass marks the tiles of bodies not at rest. */
static void solve_impl(const orc_params *p, int n, const float *in, float *out,
                       const float *materials, const uint32_t *counts, const uint32_t *lists,
                       int cap, float *pressures, int write_pressure, const sleep_ctx *sl) {
    float vw = (float)p->width, vh = (float)p->height;
    float eff_dt = p->dt / (float)p->substeps;
    for (int i = 0; i < n; i++) {
        const float *b = in + 4 * i;
        if (!sleep_awake(sl, i)) {
            memcpy(out + 4 * i, b, 16);
            continue;
        }
        float ppx = b[0] * vw, ppy = b[1] * vh;
        float vpx = b[2] * vw, vpy = b[3] * vh;
        float pcx = 0, pcy = 0, vcx = 0, vcy = 0, pressure = 0;
        uint32_t own = counts[i];
        for (uint32_t k = 0; k < own && (int)k < cap; k++) {
            uint32_t j = lists[(size_t)i * cap + k];
            accumulate_contact(p, (uint32_t)i, j, in, materials, counts, ppx, ppy, vpx, vpy, own,
                               eff_dt, !sleep_awake(sl, (int)j), &pcx, &pcy, &vcx, &vcy,
                               &pressure, write_pressure);
        }
        if (p->bounded) {
            float weight = 1.0f / (float)(own + 1u);
            float fr = materials[4 * i + 1];
            wall_contact(p, ppx, 1.0f, 0.0f, vpx, vpy, weight, fr, eff_dt, &pcx, &pcy, &vcx, &vcy, &pressure, write_pressure);
            wall_contact(p, vw - ppx, -1.0f, 0.0f, vpx, vpy, weight, fr, eff_dt, &pcx, &pcy, &vcx, &vcy, &pressure, write_pressure);
            wall_contact(p, ppy, 0.0f, 1.0f, vpx, vpy, weight, fr, eff_dt, &pcx, &pcy, &vcx, &vcy, &pressure, write_pressure);
            wall_contact(p, vh - ppy, 0.0f, -1.0f, vpx, vpy, weight, fr, eff_dt, &pcx, &pcy, &vcx, &vcy, &pressure, write_pressure);
        }
        ppx += pcx; ppy += pcy;
        vpx += vcx + pcx / eff_dt;
        vpy += vcy + pcy / eff_dt;
        float *o = out + 4 * i;
        if (p->bounded) {
            o[0] = clampf(ppx / vw, 0.0f, MAXIMUM_UNIT);
            o[1] = clampf(ppy / vh, 0.0f, MAXIMUM_UNIT);
        } else {
            ppx = glsl_mod(ppx, vw);
            ppy = glsl_mod(ppy, vh);
            o[0] = ppx / vw;
            o[1] = ppy / vh;
        }
        o[2] = vpx / vw;
        o[3] = vpy / vh;
        int fluid = p->liquid && materials[4 * i + 2] == 0.0f;
        if (write_pressure && sl && sl->enabled) {
            /* S/collision_cache_solve.comp:449-471: a body that moved this
             * step, or that nothing holds, keeps its tile awake. */
            float radius = 0.5f * p->diameter, slack = 0.05f * p->diameter;
            int held = own > 0u || (p->bounded && (ppx < radius + slack || ppy < radius + slack ||
                                                   ppx > vw - radius - slack ||
                                                   ppy > vh - radius - slack));
            if (fluid) held = 1;
            float dx = vpx * p->dt, dy = vpy * p->dt;
            float settling = 0.005f * p->diameter;
            float d2 = dx * dx + dy * dy, m2 = pcx * pcx + pcy * pcy;
            if (!held || d2 >= sl->disp2 || m2 > settling * settling) {
                sl->tile_active[sleep_tile(sl, o[0], o[1])] = sl->stamp + 1u;
                float mark = 0.5f * (float)SLEEP_TILE_PIXELS;
                if (d2 >= mark * mark)
                    sl->tile_active[sleep_tile(sl, o[0] + dx / vw, o[1] + dy / vh)] = sl->stamp + 1u;
            }
        }
        if (write_pressure && !fluid) pressures[i] = pressure;
    }
}

void orc_solve(const orc_params *p, int n, const float *in, float *out, const float *materials,
               const uint32_t *counts, const uint32_t *lists, int cap, float *pressures,
               int write_pressure) {
    solve_impl(p, n, in, out, materials, counts, lists, cap, pressures, write_pressure, NULL);
}

/* One pass: drift, contacts, solve. counts (nullable) persists between
 * passes: a sleeping body keeps its last count there
 * (S/collision_cache_count.comp:147-160). sl (nullable): sleeping tiles. */
static void collide_pass_impl(const orc_params *p, int n, float *bodies, const float *accel,
                              const float *materials, float *pressures, float kick_dt,
                              float drift_dt, int write_pressure, uint32_t *counts_out,
                              uint32_t *persist, const sleep_ctx *sl) {
    if (sl && sl->sleeping) {
        /* S/collision_cache_drift.comp:60-91: the kick pass tells every
         * body's state for the step; a sleeping body rests, its velocity
         * dropped, and neither kicks nor drifts. */
        for (int i = 0; i < n; i++) {
            float *b = bodies + 4 * i;
            int awake;
            if (kick_dt != 0.0f) {
                int fluid = p->liquid && materials[4 * i + 2] == 0.0f;
                awake = fluid || sl->tile_awake[sleep_tile(sl, b[0], b[1])] == sl->stamp;
                sl->body_awake[i] = awake ? sl->stamp : 0u;
                if (!awake) { b[2] = 0.0f; b[3] = 0.0f; }
            } else {
                awake = sl->body_awake[i] == sl->stamp;
            }
            if (!awake || (kick_dt == 0.0f && drift_dt == 0.0f)) continue;
            if (kick_dt != 0.0f) {
                float h = 0.5f * kick_dt;
                b[2] += accel[2 * i] * h;
                b[3] += accel[2 * i + 1] * h;
            }
            b[0] += b[2] * drift_dt;
            b[1] += b[3] * drift_dt;
            if (p->bounded) {
                b[0] = clampf(b[0], 0.0f, MAXIMUM_UNIT);
                b[1] = clampf(b[1], 0.0f, MAXIMUM_UNIT);
            } else {
                b[0] = fractf(b[0]);
                b[1] = fractf(b[1]);
            }
        }
    } else {
        orc_drift(p, n, bodies, accel, kick_dt, drift_dt);
    }
    if (!p->collide) return;
    uint32_t *counts = malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
    orc_contacts(p, n, bodies, materials, counts, NULL, 0);
    int cap = 1;
    for (int i = 0; i < n; i++) if ((int)counts[i] > cap) cap = (int)counts[i];
    uint32_t *lists = malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1) * (size_t)cap);
    orc_contacts(p, n, bodies, materials, counts, lists, cap);
    if (persist) {
        for (int i = 0; i < n; i++) {
            if (sleep_awake(sl, i)) persist[i] = counts[i];
            else counts[i] = persist[i];
        }
    }
    float *out = malloc((size_t)(n > 0 ? n : 1) * 16);
    solve_impl(p, n, bodies, out, materials, counts, lists, cap, pressures, write_pressure, sl);
    memcpy(bodies, out, (size_t)n * 16);
    if (counts_out) memcpy(counts_out, counts, sizeof(uint32_t) * (size_t)n);
    free(out); free(lists); free(counts);
}

void orc_collide_pass(const orc_params *p, int n, float *bodies, const float *accel,
                      const float *materials, float *pressures, float kick_dt, float drift_dt,
                      int write_pressure, uint32_t *counts_out) {
    collide_pass_impl(p, n, bodies, accel, materials, pressures, kick_dt, drift_dt,
                      write_pressure, counts_out, NULL, NULL);
}

/* ------------------------------------------------------------------------ */
/* b: P3M gravity                                                           */

/* K/GravityPipeline.kt:1546-1547 */
int orc_mass_scale(float total_mass) {
    double v = 4000000000.0 / (double)total_mass;
    long long s = (long long)v;
    if (s < 1) s = 1;
    if (s > 65536) s = 65536;
    return (int)s;
}

/* S/twiddle_init.comp:11-25, one table for every axis length. */
void orc_twiddles(int max_axis, float *tw) {
    const float TAU = 6.28318530717958647692f;
    for (int h = 1; h < max_axis; h *= 2) {
        for (int id = 0; id < h; id++) {
            float fa = -TAU * (float)id / (float)(h * 2);
            float ia = TAU * (float)id / (float)(h * 2);
            float *t = tw + 4 * (h - 1 + id);
            t[0] = cosf(fa); t[1] = sinf(fa); t[2] = cosf(ia); t[3] = sinf(ia);
        }
    }
}

void orc_twiddles_f64(int max_axis, double *tw) {
    const double TAU = 6.283185307179586476925;
    for (int h = 1; h < max_axis; h *= 2) {
        for (int id = 0; id < h; id++) {
            double fa = -TAU * id / (double)(h * 2), ia = TAU * id / (double)(h * 2);
            double *t = tw + 4 * (h - 1 + id);
            t[0] = cos(fa); t[1] = sin(fa); t[2] = cos(ia); t[3] = sin(ia);
        }
    }
}

static uint32_t bit_reverse(uint32_t v, int bits) {
    uint32_t r = 0;
    for (int b = 0; b < bits; b++) r |= ((v >> b) & 1u) << (bits - 1 - b);
    return r;
}

/* S/fft_shared.comp:26-81: bit-reversed load, radix-2 DIT stages. */
#define FFT_LINE(T, NAME)                                                              \
    static void NAME(T *line, int L, const T *tw, int inverse) {                       \
        int bits = 0;                                                                  \
        while ((1 << bits) < L) bits++;                                                \
        T *tmp = malloc(sizeof(T) * 2 * (size_t)L);                                    \
        for (int c = 0; c < L; c++) {                                                  \
            uint32_t s = bit_reverse((uint32_t)c, bits);                               \
            tmp[2 * c] = line[2 * s];                                                  \
            tmp[2 * c + 1] = line[2 * s + 1];                                          \
        }                                                                              \
        for (int sw = 2; sw <= L; sw <<= 1) {                                          \
            int half = sw >> 1;                                                        \
            for (int b = 0; b < L / 2; b++) {                                          \
                int within = b & (half - 1);                                           \
                int lower = (b / half) * sw + within;                                  \
                const T *t = tw + 4 * (half - 1 + within) + (inverse ? 2 : 0);         \
                T ur = tmp[2 * (lower + half)], ui = tmp[2 * (lower + half) + 1];      \
                T xr = t[0] * ur - t[1] * ui;                                          \
                T xi = t[0] * ui + t[1] * ur;                                          \
                T lr = tmp[2 * lower], li = tmp[2 * lower + 1];                        \
                tmp[2 * lower] = lr + xr;                                              \
                tmp[2 * lower + 1] = li + xi;                                          \
                tmp[2 * (lower + half)] = lr - xr;                                     \
                tmp[2 * (lower + half) + 1] = li - xi;                                 \
            }                                                                          \
        }                                                                              \
        memcpy(line, tmp, sizeof(T) * 2 * (size_t)L);                                  \
        free(tmp);                                                                     \
    }
FFT_LINE(float, fft_line_f32)
FFT_LINE(double, fft_line_f64)

#define FFT_2D(T, NAME, LINE)                                                          \
    void NAME(T *values, int w, int h, const T *tw, int inverse) {                     \
        T *line = malloc(sizeof(T) * 2 * (size_t)(w > h ? w : h));                     \
        for (int y = 0; y < h; y++) LINE(values + (size_t)2 * y * w, w, tw, inverse);  \
        for (int x = 0; x < w; x++) {                                                  \
            for (int y = 0; y < h; y++) {                                              \
                line[2 * y] = values[2 * ((size_t)y * w + x)];                         \
                line[2 * y + 1] = values[2 * ((size_t)y * w + x) + 1];                 \
            }                                                                          \
            LINE(line, h, tw, inverse);                                                \
            for (int y = 0; y < h; y++) {                                              \
                values[2 * ((size_t)y * w + x)] = line[2 * y];                         \
                values[2 * ((size_t)y * w + x) + 1] = line[2 * y + 1];                 \
            }                                                                          \
        }                                                                              \
        free(line);                                                                    \
    }
FFT_2D(float, orc_fft2d, fft_line_f32)
FFT_2D(double, orc_fft2d_f64, fft_line_f64)

/* S/kernel_init.comp:16-46 */
void orc_kernel_init(int mw, int mh, int pw, int ph, float aspect, float G, float total_mass,
                     float sigma_cells, float *kx, float *ky) {
    for (int id = 0; id < mw * mh; id++) {
        int rx = id % mw, ry = id / mw;
        int ox = rx <= mw / 2 ? rx : rx - mw;
        int oy = ry <= mh / 2 ? ry : ry - mh;
        float dx = (float)ox * aspect / (float)pw;
        float dy = (float)oy / (float)ph;
        float r2 = dx * dx + dy * dy;
        kx[2 * id] = kx[2 * id + 1] = ky[2 * id] = ky[2 * id + 1] = 0.0f;
        if (r2 == 0.0f) continue;
        float sigma = sigma_cells / (float)ph;
        float lw = 1.0f - expf(-r2 / (2.0f * sigma * sigma));
        float dirx = rx * 2 == mw ? 0.0f : dx;
        float diry = ry * 2 == mh ? 0.0f : dy;
        float scale = -G * lw / (total_mass * r2);
        kx[2 * id] = scale * dirx / aspect;
        ky[2 * id] = scale * diry;
    }
}

static float sincf(float v) { return fabsf(v) < 1e-6f ? 1.0f : sinf(v) / v; }

/* S/convolve.comp:24-47 */
void orc_convolve(int mw, int mh, float window_floor, const float *rho, const float *kx,
                  const float *ky, float *fx, float *fy) {
    const float PI = 3.14159265358979323846f;
    for (int id = 0; id < mw * mh; id++) {
        int rx = id % mw, ry = id / mw;
        int ikx = rx <= mw / 2 ? rx : rx - mw;
        int iky = ry <= mh / 2 ? ry : ry - mh;
        float sx = sincf(PI * (float)ikx / (float)mw);
        float sy = sincf(PI * (float)iky / (float)mh);
        float aw = sx * sx * sy * sy;
        float p2 = aw * aw;
        float deconv = 1.0f / (p2 > window_floor ? p2 : window_floor);
        float inv = 1.0f / (float)(mw * mh);
        float scale = deconv * inv;
        float rr = rho[2 * id], ri = rho[2 * id + 1];
        const float *a = kx + 2 * id, *b = ky + 2 * id;
        fx[2 * id] = (rr * a[0] - ri * a[1]) * scale;
        fx[2 * id + 1] = (rr * a[1] + ri * a[0]) * scale;
        fy[2 * id] = (rr * b[0] - ri * b[1]) * scale;
        fy[2 * id + 1] = (rr * b[1] + ri * b[0]) * scale;
    }
}

static int wrap_axis(int c, int size) {
    if (c < 0) c += size; else if (c >= size) c -= size;
    return c;
}

/* S/build_grid.comp:49-76 (fixed-point CIC; the linked lists are replaced by
 * per-cell tables, see near_field below). round() is round-half-even. */
void orc_deposit(int n, const float *bodies, const float *materials, int mw, int mh, int pw,
                 int ph, int mass_scale, uint32_t *density, int32_t *cell_count,
                 uint32_t *cell_mass) {
    memset(density, 0, sizeof(uint32_t) * (size_t)mw * mh);
    memset(cell_count, 0, sizeof(int32_t) * (size_t)pw * ph);
    memset(cell_mass, 0, sizeof(uint32_t) * (size_t)pw * ph);
    float scale = (float)mass_scale;
    for (int i = 0; i < n; i++) {
        float px = bodies[4 * i], py = bodies[4 * i + 1];
        float mass = materials[4 * i];
        float gx = px * (float)pw - 0.5f, gy = py * (float)ph - 0.5f;
        float flx = floorf(gx), fly = floorf(gy);
        int bx = (int)flx, by = (int)fly;
        float fx = gx - flx, fy = gy - fly;
        float w[4] = {(1.0f - fx) * (1.0f - fy), fx * (1.0f - fy), (1.0f - fx) * fy, fx * fy};
        int ox[4] = {0, 1, 0, 1}, oy[4] = {0, 0, 1, 1};
        for (int k = 0; k < 4; k++) {
            float wk = w[k] * mass;
            uint32_t q = (uint32_t)rintf(wk * scale);
            int mx = wrap_axis(bx + ox[k], mw), my = wrap_axis(by + oy[k], mh);
            density[(size_t)my * mw + mx] += q;
        }
        int hx = (int)floorf(px * (float)pw), hy = (int)floorf(py * (float)ph);
        int home = wrap_axis(hx, pw) + wrap_axis(hy, ph) * pw;
        cell_count[home] += 1;
        cell_mass[home] += (uint32_t)rintf(mass * 256.0f);
    }
}

/* K/GravityPipeline.kt:153-169 */
void orc_neighbor_radius(const orc_params *p, int *rx, int *ry) {
    float aspect = (float)p->width / (float)p->height;
    double x = ceil((double)(p->near_sigma * p->softening_cells * (float)p->grid_w /
                             ((float)p->grid_h * aspect)));
    double y = ceil((double)(p->softening_cells * p->near_sigma));
    int ix = (int)x, iy = (int)y;
    *rx = clampi(ix, p->near_min, p->near_max);
    *ry = clampi(iy, p->near_min, p->near_max);
}

/* ------------------------------------------------------------------------ */
/* The simulation (K/GravityPipeline.kt:445-494)                            */

struct orc_sim {
    orc_params p;
    int n;
    float *bodies, *accel, *mats, *press;
    int mw, mh, pw, ph;
    float total_mass;
    float *tw, *kx, *ky, *fft, *fx, *fy;
    uint32_t *density;
    int32_t *cell_count, *cell_start, *cell_list;
    uint32_t *cell_mass;
    long steps;
    uint32_t *last_dest;
    int kernel_valid;
    uint32_t *counts;         /* contact counts, kept between passes and steps */
    sleep_ctx sleep;
    int wake_steps;
};

static void sim_init_kernel(orc_sim *s) {
    float aspect = (float)s->p.width / (float)s->p.height;
    orc_kernel_init(s->mw, s->mh, s->pw, s->ph, aspect, s->p.gravity, s->total_mass,
                    s->p.softening_cells, s->kx, s->ky);
    orc_fft2d(s->kx, s->mw, s->mh, s->tw, 0);
    orc_fft2d(s->ky, s->mw, s->mh, s->tw, 0);
    s->kernel_valid = 1;
}

orc_sim *orc_sim_create(const orc_params *p, int n, const float *bodies, const float *materials) {
    orc_sim *s = calloc(1, sizeof(orc_sim));
    s->p = *p;
    s->n = n;
    size_t nn = (size_t)(n > 0 ? n : 1);
    s->bodies = malloc(nn * 16); s->accel = calloc(nn * 2, 4);
    s->mats = malloc(nn * 16); s->press = calloc(nn, 4);
    memcpy(s->bodies, bodies, (size_t)n * 16);
    if (materials) memcpy(s->mats, materials, (size_t)n * 16);
    else for (int i = 0; i < n; i++) { s->mats[4*i] = 1; s->mats[4*i+1] = 0; s->mats[4*i+2] = 0; s->mats[4*i+3] = 0; }
    s->pw = p->grid_w; s->ph = p->grid_h;
    s->mw = p->bounded ? 2 * s->pw : s->pw;
    s->mh = p->bounded ? 2 * s->ph : s->ph;
    int max_axis = s->mw > s->mh ? s->mw : s->mh;
    s->tw = malloc(sizeof(float) * 4 * (size_t)max_axis);
    orc_twiddles(max_axis, s->tw);
    size_t m = (size_t)s->mw * s->mh;
    s->kx = malloc(8 * m); s->ky = malloc(8 * m); s->fft = malloc(8 * m);
    s->fx = malloc(8 * m); s->fy = malloc(8 * m);
    s->density = malloc(4 * m);
    size_t cells = (size_t)s->pw * s->ph;
    s->cell_count = malloc(4 * cells); s->cell_start = malloc(4 * (cells + 1));
    s->cell_mass = malloc(4 * cells); s->cell_list = malloc(4 * nn);
    s->last_dest = malloc(4 * nn);
    for (int i = 0; i < n; i++) s->last_dest[i] = (uint32_t)i;
    s->total_mass = (float)(n > 1 ? n : 1);
    s->counts = calloc(nn, 4);
    /* K/SleepTiles.kt:37-60: 32 px tiles, stamp 1, everything awake for
     * the first two steps. */
    s->sleep.enabled = p->sleep && p->collide;
    s->sleep.stamp = 1u;
    s->sleep.tiles_x = (p->width + SLEEP_TILE_PIXELS - 1) / SLEEP_TILE_PIXELS;
    s->sleep.tiles_y = (p->height + SLEEP_TILE_PIXELS - 1) / SLEEP_TILE_PIXELS;
    size_t tiles = (size_t)s->sleep.tiles_x * s->sleep.tiles_y;
    s->sleep.tile_awake = calloc(tiles, 4);
    s->sleep.tile_active = calloc(tiles, 4);
    s->sleep.body_awake = calloc(nn, 4);
    float disp = p->sleep_speed * p->dt;
    s->sleep.disp2 = disp * disp;
    s->wake_steps = SLEEP_WAKE_STEPS;
    return s;
}

void orc_sim_wake(orc_sim *s) { s->wake_steps = SLEEP_WAKE_STEPS; }
uint32_t orc_sim_sleep_stamp(const orc_sim *s) { return s->sleep.stamp; }
void orc_sim_body_awake(const orc_sim *s, uint32_t *awake) {
    memcpy(awake, s->sleep.body_awake, (size_t)s->n * 4);
}

/* S/sleep_tiles.comp:10-28 then K/SleepTiles.kt:91-100 endStep: the tiles
 * awake in the next step are those marked active or next to one. */
static void sleep_advance(orc_sim *s) {
    sleep_ctx *sl = &s->sleep;
    if (!sl->enabled) return;
    uint32_t next = sl->stamp + 1u;
    int tx = sl->tiles_x, ty = sl->tiles_y;
    for (int index = 0; index < tx * ty; index++) {
        int x0 = index % tx, y0 = index / tx, awake = 0;
        for (int dy = -1; dy <= 1; dy++) {
            for (int dx = -1; dx <= 1; dx++) {
                int x = x0 + dx, y = y0 + dy;
                if (s->p.bounded) {
                    if (x < 0 || y < 0 || x >= tx || y >= ty) continue;
                } else {
                    x = (x + tx) % tx;
                    y = (y + ty) % ty;
                }
                if (sl->tile_active[x + y * tx] == next) awake = 1;
            }
        }
        if (awake) sl->tile_awake[index] = next;
    }
    sl->stamp++;
    if (sl->stamp == 0u) sl->stamp = 1u;
    if (s->wake_steps > 0) s->wake_steps--;
}

void orc_sim_destroy(orc_sim *s) {
    if (!s) return;
    free(s->bodies); free(s->accel); free(s->mats); free(s->press);
    free(s->tw); free(s->kx); free(s->ky); free(s->fft); free(s->fx); free(s->fy);
    free(s->density); free(s->cell_count); free(s->cell_start); free(s->cell_mass);
    free(s->cell_list); free(s->last_dest); free(s->counts);
    free(s->sleep.tile_awake); free(s->sleep.tile_active); free(s->sleep.body_awake); free(s);
}

void orc_sim_set_bodies(orc_sim *s, const float *bodies) { memcpy(s->bodies, bodies, (size_t)s->n * 16); }
void orc_sim_set_accel(orc_sim *s, const float *accel) { memcpy(s->accel, accel, (size_t)s->n * 8); }

void orc_sim_get(const orc_sim *s, float *bodies, float *accel, float *pressures, float *materials) {
    if (bodies) memcpy(bodies, s->bodies, (size_t)s->n * 16);
    if (accel) memcpy(accel, s->accel, (size_t)s->n * 8);
    if (pressures) memcpy(pressures, s->press, (size_t)s->n * 4);
    if (materials) memcpy(materials, s->mats, (size_t)s->n * 16);
}

long orc_sim_steps(const orc_sim *s) { return s->steps; }
void orc_sim_last_sort_dest(const orc_sim *s, uint32_t *dest) { memcpy(dest, s->last_dest, (size_t)s->n * 4); }

/* K/GravityPipeline.kt:1326-1331 */
static void sim_normalize(orc_sim *s) {
    float mass = (float)(s->n > 1 ? s->n : 1);
    if (mass == s->total_mass && s->kernel_valid) return;
    s->total_mass = mass;
    sim_init_kernel(s);
}

/* K/GravityPipeline.kt:907-973 */
void orc_sim_build_force(orc_sim *s) {
    if (!s->kernel_valid) sim_init_kernel(s);
    int scale = orc_mass_scale(s->total_mass);
    orc_deposit(s->n, s->bodies, s->mats, s->mw, s->mh, s->pw, s->ph, scale, s->density,
                s->cell_count, s->cell_mass);
    /* Cell lists in descending body index (the canonical list order). */
    size_t cells = (size_t)s->pw * s->ph;
    s->cell_start[0] = 0;
    for (size_t c = 0; c < cells; c++) s->cell_start[c + 1] = s->cell_start[c] + s->cell_count[c];
    int32_t *fill = malloc(4 * cells);
    for (size_t c = 0; c < cells; c++) fill[c] = s->cell_start[c + 1];
    for (int i = 0; i < s->n; i++) {
        float px = s->bodies[4 * i], py = s->bodies[4 * i + 1];
        int hx = (int)floorf(px * (float)s->pw), hy = (int)floorf(py * (float)s->ph);
        int home = wrap_axis(hx, s->pw) + wrap_axis(hy, s->ph) * s->pw;
        s->cell_list[--fill[home]] = i;
    }
    free(fill);
    /* S/fft_init.comp:11-17 */
    float inv = 1.0f / (float)scale;
    size_t m = (size_t)s->mw * s->mh;
    for (size_t i = 0; i < m; i++) { s->fft[2 * i] = (float)s->density[i] * inv; s->fft[2 * i + 1] = 0.0f; }
    orc_fft2d(s->fft, s->mw, s->mh, s->tw, 0);
    orc_convolve(s->mw, s->mh, s->p.window_floor, s->fft, s->kx, s->ky, s->fx, s->fy);
    orc_fft2d(s->fx, s->mw, s->mh, s->tw, 1);
    orc_fft2d(s->fy, s->mw, s->mh, s->tw, 1);
}

void orc_sim_forces(const orc_sim *s, float *fx, float *fy) {
    size_t m = (size_t)s->mw * s->mh;
    for (size_t i = 0; i < m; i++) { fx[i] = s->fx[2 * i]; fy[i] = s->fy[2 * i]; }
}

void orc_sim_density_spectrum(const orc_sim *s, float *spectrum) {
    memcpy(spectrum, s->fft, 8 * (size_t)s->mw * s->mh);
}

/* S/kick.comp:66-80 */
static void local_acceleration(const orc_sim *s, float sx, float sy, float tx, float ty,
                               float mass, float *ax, float *ay) {
    float aspect = (float)s->p.width / (float)s->p.height;
    float dx = sx - tx, dy = sy - ty;
    if (!s->p.bounded) { dx -= floorf(dx + 0.5f); dy -= floorf(dy + 0.5f); }
    float px = dx * aspect, py = dy * 1.0f;
    float r2 = px * px + py * py;
    float sigma = s->p.softening_cells / (float)s->ph;
    float lw = expf(-r2 / (2.0f * sigma * sigma));
    float cw = aspect / (float)s->pw, ch = 1.0f / (float)s->ph;
    float md = s->p.min_dist_frac * (cw < ch ? cw : ch);
    float mm = md * md;
    float scale = s->p.gravity * mass * lw / (s->total_mass * (r2 > mm ? r2 : mm));
    *ax += scale * px / aspect;
    *ay += scale * py / 1.0f;
}

static float mixf(float x, float y, float a) { return x * (1.0f - a) + y * a; }

/* S/kick.comp:82-103 */
static void interpolate_force(const orc_sim *s, float px, float py, float *ax, float *ay) {
    float gx = px * (float)s->pw - 0.5f, gy = py * (float)s->ph - 0.5f;
    float flx = floorf(gx), fly = floorf(gy);
    int bx = (int)flx, by = (int)fly;
    float fx = gx - flx, fy = gy - fly;
    float vx[4], vy[4];
    int ox[4] = {0, 1, 0, 1}, oy[4] = {0, 0, 1, 1};
    for (int k = 0; k < 4; k++) {
        int mx = wrap_axis(bx + ox[k], s->mw), my = wrap_axis(by + oy[k], s->mh);
        size_t idx = (size_t)my * s->mw + mx;
        vx[k] = s->fx[2 * idx];
        vy[k] = s->fy[2 * idx];
    }
    *ax = mixf(mixf(vx[0], vx[1], fx), mixf(vx[2], vx[3], fx), fy);
    *ay = mixf(mixf(vy[0], vy[1], fx), mixf(vy[2], vy[3], fx), fy);
}

/* S/kick.comp:105-156 (sleep off) */
void orc_sim_kick(orc_sim *s, int initial) {
    float fieldx = s->p.field_x / (float)s->p.width, fieldy = s->p.field_y / (float)s->p.height;
    float half = 0.5f * s->p.dt;
    int rx, ry;
    orc_neighbor_radius(&s->p, &rx, &ry);
    for (int i = 0; i < s->n; i++) {
        float *b = s->bodies + 4 * i;
        /* S/kick.comp:109-110: a sleeping body is not kicked. */
        if (!initial && !sleep_awake(&s->sleep, i)) continue;
        if (!s->p.mutual_gravity) {
            if (!initial) { b[2] += fieldx * half; b[3] += fieldy * half; }
            s->accel[2 * i] = fieldx; s->accel[2 * i + 1] = fieldy;
            continue;
        }
        float ax, ay;
        interpolate_force(s, b[0], b[1], &ax, &ay);
        ax += fieldx; ay += fieldy;
        int hx = (int)floorf(b[0] * (float)s->pw), hy = (int)floorf(b[1] * (float)s->ph);
        for (int dy = -ry; dy <= ry; dy++) {
            for (int dx = -rx; dx <= rx; dx++) {
                int cx = hx + dx, cy = hy + dy;
                if (s->p.bounded && (cx < 0 || cy < 0 || cx >= s->pw || cy >= s->ph)) continue;
                int wx = wrap_axis(cx, s->pw), wy = wrap_axis(cy, s->ph);
                int index = wx + wy * s->pw;
                int count = s->cell_count[index];
                int home = dx == 0 && dy == 0;
                if (!home && count > s->p.exact_limit) {
                    float cxn = ((float)wx + 0.5f) / (float)s->pw;
                    float cyn = ((float)wy + 0.5f) / (float)s->ph;
                    float agg = (float)s->cell_mass[index] / 256.0f;
                    local_acceleration(s, cxn, cyn, b[0], b[1], agg, &ax, &ay);
                } else {
                    int limit = home ? s->p.home_limit : s->p.exact_limit;
                    int visited = 0;
                    for (int t = s->cell_start[index]; t < s->cell_start[index + 1] && visited < limit; t++) {
                        int other = s->cell_list[t];
                        if (other != i)
                            local_acceleration(s, s->bodies[4 * other], s->bodies[4 * other + 1], b[0],
                                               b[1], s->mats[4 * other], &ax, &ay);
                        visited++;
                    }
                }
            }
        }
        if (!initial) { b[2] += ax * half; b[3] += ay * half; }
        s->accel[2 * i] = ax; s->accel[2 * i + 1] = ay;
    }
}

/* ---- the trajectory preview (S/trajectory.comp) -------------------------- */

/* S/trajectory.comp:77-98: localAcceleration, the test particle being no body */
static void trajectory_local(const orc_sim *s, float sx, float sy, float tx, float ty, float mass,
                             float *ax, float *ay) {
    float aspect = (float)s->p.width / (float)s->p.height;
    float dx = sx - tx, dy = sy - ty;
    if (!s->p.bounded) { dx -= floorf(dx + 0.5f); dy -= floorf(dy + 0.5f); }
    float px = dx * aspect, py = dy * 1.0f;
    float r2 = px * px + py * py;
    float sigma = s->p.softening_cells / (float)s->ph;
    float lw = expf(-r2 / (2.0f * sigma * sigma));
    float cw = aspect / (float)s->pw, ch = 1.0f / (float)s->ph;
    float md = s->p.min_dist_frac * (cw < ch ? cw : ch);
    float scale = s->p.gravity * mass * lw / (s->total_mass * (r2 > md * md ? r2 : md * md));
    *ax += scale * px / aspect;
    *ay += scale * py / 1.0f;
}

/* S/trajectory.comp:124-168: accelerationAt */
static void trajectory_acceleration(const orc_sim *s, float x, float y, float *ax, float *ay) {
    float fieldx = s->p.field_x / (float)s->p.width, fieldy = s->p.field_y / (float)s->p.height;
    if (!s->p.mutual_gravity) { *ax = fieldx; *ay = fieldy; return; }
    interpolate_force(s, x, y, ax, ay);
    *ax += fieldx;
    *ay += fieldy;
    int rx, ry;
    orc_neighbor_radius(&s->p, &rx, &ry);
    int hx = (int)floorf(x * (float)s->pw), hy = (int)floorf(y * (float)s->ph);
    for (int dy = -ry; dy <= ry; dy++) {
        for (int dx = -rx; dx <= rx; dx++) {
            int cx = hx + dx, cy = hy + dy;
            if (s->p.bounded && (cx < 0 || cy < 0 || cx >= s->pw || cy >= s->ph)) continue;
            int wx = wrap_axis(cx, s->pw), wy = wrap_axis(cy, s->ph);
            int index = wx + wy * s->pw;
            int count = s->cell_count[index];
            int home = dx == 0 && dy == 0;
            if (!home && count > s->p.exact_limit) {
                float cxn = ((float)wx + 0.5f) / (float)s->pw;
                float cyn = ((float)wy + 0.5f) / (float)s->ph;
                trajectory_local(s, cxn, cyn, x, y, (float)s->cell_mass[index] / 256.0f, ax, ay);
            } else {
                int limit = home ? s->p.home_limit : s->p.exact_limit;
                int visited = 0;
                for (int t = s->cell_start[index]; t < s->cell_start[index + 1] && visited < limit; t++) {
                    int other = s->cell_list[t];
                    trajectory_local(s, s->bodies[4 * other], s->bodies[4 * other + 1], x, y,
                                     s->mats[4 * other], ax, ay);
                    visited++;
                }
            }
        }
    }
}

/* S/trajectory.comp:170-186: integrationDt */
static float trajectory_dt(const orc_sim *s, const orc_trajectory_params *t, float vx, float vy,
                           float ax, float ay) {
    float W = (float)s->p.width, H = (float)s->p.height;
    float sx = vx * W, sy = vy * H, gx = ax * W, gy = ay * H;
    float speed = sqrtf(sx * sx + sy * sy);
    float accel = sqrtf(gx * gx + gy * gy);
    float r;
    if (accel > 1e-5f)
        r = (sqrtf(speed * speed + 2.0f * accel * t->segment_px) - speed) / accel;
    else
        r = t->segment_px / (speed > 1e-5f ? speed : 1e-5f);
    return clampf(r, s->p.dt, t->maximum_dt);
}

/* S/trajectory.comp:188-259: emitPeriodicStep */
static void trajectory_emit(float *out, int step, float sx, float sy, float dx, float dy,
                            float start_distance, float step_distance, float start_dash,
                            float step_dash) {
    float *v = out + (size_t)step * 6 * 4;
    for (int k = 0; k < 6; k++) { v[4 * k] = 0.0f; v[4 * k + 1] = 0.0f; v[4 * k + 2] = 0.0f; v[4 * k + 3] = -1.0f; }
    float ex = sx + dx, ey = sy + dy;
    float cuts[4];
    int cut_count = 2;
    cuts[0] = 0.0f;
    cuts[1] = 1.0f;
    if (ex < 0.0f) cuts[cut_count++] = -sx / dx;
    else if (ex >= 1.0f) cuts[cut_count++] = (1.0f - sx) / dx;
    if (ey < 0.0f) cuts[cut_count++] = -sy / dy;
    else if (ey >= 1.0f) cuts[cut_count++] = (1.0f - sy) / dy;
    for (int index = 1; index < cut_count; index++) {
        float value = clampf(cuts[index], 0.0f, 1.0f);
        int insertion = index;
        while (insertion > 0 && cuts[insertion - 1] > value) {
            cuts[insertion] = cuts[insertion - 1];
            insertion--;
        }
        cuts[insertion] = value;
    }
    float unique[4];
    int unique_count = 0;
    for (int index = 0; index < cut_count; index++)
        if (unique_count == 0 || fabsf(cuts[index] - unique[unique_count - 1]) > 1e-6f)
            unique[unique_count++] = cuts[index];
    for (int piece = 0; piece < unique_count - 1; piece++) {
        float f1 = unique[piece], f2 = unique[piece + 1];
        float mid = 0.5f * (f1 + f2);
        float tx = floorf(sx + dx * mid), ty = floorf(sy + dy * mid);
        float *a = v + 4 * (piece * 2);
        a[0] = clampf(sx + dx * f1 - tx, 0.0f, 1.0f);
        a[1] = clampf(sy + dy * f1 - ty, 0.0f, 1.0f);
        a[2] = start_dash + step_dash * f1;
        a[3] = start_distance + step_distance * f1;
        a[4] = clampf(sx + dx * f2 - tx, 0.0f, 1.0f);
        a[5] = clampf(sy + dy * f2 - ty, 0.0f, 1.0f);
        a[6] = start_dash + step_dash * f2;
        a[7] = start_distance + step_distance * f2;
    }
}

/* S/trajectory.comp:261-329 with the uniforms of K/GravityPipeline.kt:514-598 */
void orc_sim_trajectory(const orc_sim *s, float origin_x, float origin_y, float vx_px, float vy_px,
                        const orc_trajectory_params *t, float *out) {
    float W = (float)s->p.width, H = (float)s->p.height;
    /* placeX / placeY (K/GravityPipeline.kt:773-779) */
    float ox, oy;
    if (s->p.bounded) {
        ox = clampf(origin_x, 0.0f, W * MAXIMUM_UNIT);
        oy = clampf(origin_y, 0.0f, H * MAXIMUM_UNIT);
    } else {
        ox = origin_x - floorf(origin_x / W) * W;
        oy = origin_y - floorf(origin_y / H) * H;
    }
    float max_path = t->path_diagonals * sqrtf(W * W + H * H);
    float x = ox / W, y = oy / H;
    float vx = vx_px / W, vy = vy_px / H;
    float ax, ay;
    trajectory_acceleration(s, x, y, &ax, &ay);
    float path = 0.0f, dash = 0.0f;
    int finished = 0;
    for (int step = 0; step < t->steps; step++) {
        if (finished) {
            float *v = out + (size_t)step * 6 * 4;
            for (int k = 0; k < 6; k++) { v[4 * k] = 0.0f; v[4 * k + 1] = 0.0f; v[4 * k + 2] = 0.0f; v[4 * k + 3] = -1.0f; }
            continue;
        }
        float dt = trajectory_dt(s, t, vx, vy, ax, ay);
        float hvx = vx + ax * (0.5f * dt), hvy = vy + ay * (0.5f * dt);
        float dx = hvx * dt, dy = hvy * dt;
        if (s->p.bounded) {
            /* walls act like body contacts: the normal motion stops, and
               rebounds by the restitution coefficient */
            float nx = x + dx, ny = y + dy;
            float cx = clampf(nx, 0.0f, MAXIMUM_UNIT), cy = clampf(ny, 0.0f, MAXIMUM_UNIT);
            /* mix(v, -e v, hit) as GLSL forms it */
            float hx = nx != cx ? 1.0f : 0.0f, hy = ny != cy ? 1.0f : 0.0f;
            hvx = hvx * (1.0f - hx) + (-s->p.restitution * hvx) * hx;
            hvy = hvy * (1.0f - hy) + (-s->p.restitution * hvy) * hy;
            dx = cx - x;
            dy = cy - y;
        }
        float sdx = dx * W, sdy = dy * H;
        float step_distance = sqrtf(sdx * sdx + sdy * sdy);
        float remaining = max_path - path;
        if (step_distance > remaining) {
            float k = remaining / step_distance;
            dx *= k;
            dy *= k;
            step_distance = remaining;
            finished = 1;
        }
        float speed_multiplier = 1.0f + sqrtf(hvx * hvx + hvy * hvy);
        float dash_length = t->minimum_dash_px * speed_multiplier;
        float step_dash = step_distance * t->minimum_dash_px / dash_length;
        trajectory_emit(out, step, x, y, dx, dy, path, step_distance, dash, step_dash);
        path += step_distance;
        dash += step_dash;
        if (s->p.bounded) {
            x = x + dx;
            y = y + dy;
        } else {
            x = x + dx - floorf(x + dx);
            y = y + dy - floorf(y + dy);
        }
        if (path >= max_path - 1e-4f) {
            finished = 1;
            continue;
        }
        trajectory_acceleration(s, x, y, &ax, &ay);
        vx = hvx + ax * (0.5f * dt);
        vy = hvy + ay * (0.5f * dt);
    }
    float inv = 1.0f / (path > 1e-6f ? path : 1e-6f);
    for (size_t k = 0; k < (size_t)t->steps * 6; k++)
        if (out[4 * k + 3] >= 0.0f) out[4 * k + 3] = clampf(out[4 * k + 3] * inv, 0.0f, 1.0f);
}

/* K/GravityPipeline.kt:366-370 */
void orc_sim_initialize_acceleration(orc_sim *s) {
    sim_normalize(s);
    if (s->p.mutual_gravity) orc_sim_build_force(s);
    orc_sim_kick(s, 1);
}

/* K/GravityPipeline.kt:445-494 (sleep off, no blocks, non-adaptive) */
void orc_sim_step(orc_sim *s) {
    sim_normalize(s);
    if (s->p.sort_interval > 0 && (s->steps++ % s->p.sort_interval) == 0)
        orc_spatial_sort(s->n, s->p.width, s->p.height, s->bodies, s->accel, s->mats, s->last_dest);
    else if (s->p.sort_interval <= 0)
        s->steps++;
    /* Sleep needs the contact passes: their final pass marks the tiles. */
    s->sleep.sleeping = s->p.sleep && s->p.collide && s->wake_steps == 0;
    float sub = s->p.dt / (float)s->p.substeps;
    for (int k = 0; k < s->p.substeps; k++)
        collide_pass_impl(&s->p, s->n, s->bodies, s->accel, s->mats, s->press,
                          k == 0 ? s->p.dt : 0.0f, sub, 0, NULL, s->counts, &s->sleep);
    if (s->p.liquid) orc_sim_resolve_liquid(s);
    if (s->p.mutual_gravity) orc_sim_build_force(s);
    orc_sim_kick(s, 0);
    if (s->p.collide)
        collide_pass_impl(&s->p, s->n, s->bodies, s->accel, s->mats, s->press, 0.0f, 0.0f, 1,
                          NULL, s->counts, &s->sleep);
    sleep_advance(s);
}

/* ------------------------------------------------------------------------ */
/* d: position-based liquid (K/LiquidSolver.kt, K/LiquidKernel.kt,          */
/*    S/liquid_*.comp, S/liquid_kernel.glsl)                                 */

/* K/LiquidKernel.kt:14-60 */
void orc_liquid_constants(float diameter, float radius_diameters, float *h_out, float *poly6_out,
                          float *spiky_out, float *rest_out, float *relax_out) {
    const double PI = 3.141592653589793;
    float h = diameter * radius_diameters;
    double hd = (double)h;
    double h5 = hd * hd * hd * hd * hd, h8 = h5 * hd * hd * hd;
    float poly6 = (float)(4.0 / (PI * h8));
    float spiky = (float)(-30.0 / (PI * h5));
    double d = (double)diameter;
    double row_step = d * sqrt(3.0) * 0.5;
    int rows = (int)ceil(h / row_step) + 1;
    int cols = (int)ceil(h / d) + 1;
    double rest = 0.0;
    for (int r = -rows; r <= rows; r++) {
        double shift = (r & 1) == 0 ? 0.0 : 0.5 * d;
        for (int c = -cols; c <= cols; c++) {
            double x = c * d + shift, y = r * row_step;
            double q = (double)h * h - (x * x + y * y);
            rest += q > 0.0 ? (double)poly6 * q * q * q : 0.0;
        }
    }
    float restf = (float)rest;
    double relax = 0.0;
    for (int r = -rows; r <= rows; r++) {
        double shift = (r & 1) == 0 ? 0.0 : 0.5 * d;
        for (int c = -cols; c <= cols; c++) {
            double x = c * d + shift, y = r * row_step;
            double dist = sqrt(x * x + y * y);
            double g = 0.0;
            if (dist > 0.0 && dist < hd) { double q = hd - dist; g = -(double)spiky * q * q; }
            g /= (double)restf;
            relax += g * g;
        }
    }
    *h_out = h; *poly6_out = poly6; *spiky_out = spiky; *rest_out = restf;
    *relax_out = (float)(0.01 * relax);
}

typedef struct {
    float h, poly6, spiky, rest, relax, row_step, body_d, pressure_scale;
    float vw, vh;
    int bounded;
} liquid_k;

static float lq_poly6(const liquid_k *k, float d2) {
    float x = k->h * k->h - d2;
    return x > 0.0f ? k->poly6 * x * x * x : 0.0f;
}

static void lq_spiky(const liquid_k *k, float dx, float dy, float dist, uint32_t first,
                     uint32_t second, float *gx, float *gy) {
    if (dist >= k->h) { *gx = 0.0f; *gy = 0.0f; return; }
    float nx, ny;
    if (dist > 1e-4f) { nx = dx / dist; ny = dy / dist; }
    else fallback_normal(first, second, &nx, &ny);
    float x = k->h - dist;
    float m = k->spiky * x * x;
    *gx = nx * m; *gy = ny * m;
}

static void lq_wall_density(const liquid_k *k, float distance, float *rho, float *deriv) {
    float rx = 0.0f, ry = 0.0f;
    float h2 = k->h * k->h;
    float delta = distance - 0.5f * k->body_d + k->row_step;
    for (int row = 0; row < 8 && delta < k->h; row++) {
        float dm = delta > 0.0f ? delta : 0.0f;
        float a2 = h2 - dm * dm;
        float a = sqrtf(a2);
        float a5 = a2 * a2 * a;
        rx += a5 * a2;
        ry += -7.0f * dm * a5;
        delta += k->row_step;
    }
    float c = k->poly6 * 32.0f / 35.0f / k->body_d;
    *rho = rx * c; *deriv = ry * c;
}

static void lq_walls(const liquid_k *k, float ppx, float ppy, float *rho, float *gx, float *gy) {
    *rho = *gx = *gy = 0.0f;
    if (!k->bounded) return;
    float lo[2], hi[2], bo[2], to[2];
    lq_wall_density(k, ppx, &lo[0], &lo[1]);
    lq_wall_density(k, k->vw - ppx, &hi[0], &hi[1]);
    lq_wall_density(k, ppy, &bo[0], &bo[1]);
    lq_wall_density(k, k->vh - ppy, &to[0], &to[1]);
    *rho = lo[0] + hi[0] + bo[0] + to[0];
    *gx = lo[1] - hi[1];
    *gy = bo[1] - to[1];
}

static void lq_delta(const liquid_k *k, const float *a, const float *b, float *dx, float *dy) {
    *dx = (a[0] - b[0]) * k->vw;
    *dy = (a[1] - b[1]) * k->vh;
    if (!k->bounded) {
        *dx -= floorf(*dx / k->vw + 0.5f) * k->vw;
        *dy -= floorf(*dy / k->vh + 0.5f) * k->vh;
    }
}

/* K/LiquidSolver.kt:107-173: grid of the step's start positions, 12 Jacobi
 * (lambda, apply) iterations, XSPH viscosity. The neighbours of a body are
 * the bodies of the 3x3 cells around its start cell (S/liquid_kernel.glsl:
 * "every pass searches around the cell a body had when the grid was
 * built"). The reference sums them in the order of its atomically built
 * lists; the canonical order here is the stencil's cell order (rows, then
 * columns, from -1 to 1), ascending body index within a cell, a cell met
 * twice on a grid narrower than three cells taken once. */
void orc_sim_resolve_liquid(orc_sim *s) {
    const orc_params *p = &s->p;
    liquid_k k;
    orc_liquid_constants(p->diameter, p->kernel_radius_diameters, &k.h, &k.poly6, &k.spiky,
                         &k.rest, &k.relax);
    k.row_step = p->diameter * sqrtf(3.0f) * 0.5f;
    k.body_d = p->diameter;
    k.pressure_scale = k.h / p->dt;
    k.vw = (float)p->width; k.vh = (float)p->height;
    k.bounded = p->bounded;
    int n = s->n;
    int gw = (int)((float)p->width / k.h); if (gw < 1) gw = 1;
    int gh = (int)((float)p->height / k.h); if (gh < 1) gh = 1;
    int *cell = malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int *start = calloc((size_t)gw * gh + 1, sizeof(int));
    int *items = malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < n; i++) {
        int cx = floor_index(s->bodies[4 * i] * (float)gw, gw);
        int cy = floor_index(s->bodies[4 * i + 1] * (float)gh, gh);
        cell[i] = cx + cy * gw;
        start[cell[i] + 1]++;
    }
    for (int c = 0; c < gw * gh; c++) start[c + 1] += start[c];
    int *fill = malloc(sizeof(int) * (size_t)gw * gh);
    memcpy(fill, start, sizeof(int) * (size_t)gw * gh);
    for (int i = 0; i < n; i++) items[fill[cell[i]]++] = i;
    free(fill);
    /* neighbour candidates per body, in the canonical order */
    int *nstart = malloc(sizeof(int) * ((size_t)n + 1));
    int cap = 64, used = 0;
    int *nb = malloc(sizeof(int) * (size_t)cap);
    for (int i = 0; i < n; i++) {
        nstart[i] = used;
        int hx = cell[i] % gw, hy = cell[i] / gw;
        int first = used;
        for (int dy = -1; dy <= 1; dy++) for (int dx = -1; dx <= 1; dx++) {
            int x = hx + dx, y = hy + dy;
            if (p->bounded) { if (x < 0 || y < 0 || x >= gw || y >= gh) continue; }
            else { x = (x + gw) % gw; y = (y + gh) % gh; }
            int c = x + y * gw;
            for (int t = start[c]; t < start[c + 1]; t++) {
                int j = items[t];
                if (j == i) continue;
                int dup = 0;
                for (int q = first; q < used; q++) dup |= nb[q] == j;
                if (dup) continue;
                if (used == cap) { cap *= 2; nb = realloc(nb, sizeof(int) * (size_t)cap); }
                nb[used++] = j;
            }
        }
    }
    nstart[n] = used;
    float *pos = malloc((size_t)(n > 0 ? n : 1) * 8), *lam = malloc((size_t)(n > 0 ? n : 1) * 4);
    float *out = malloc((size_t)(n > 0 ? n : 1) * 16), *pout = malloc((size_t)(n > 0 ? n : 1) * 8);
    for (int i = 0; i < n; i++) { pos[2 * i] = s->bodies[4 * i]; pos[2 * i + 1] = s->bodies[4 * i + 1]; }
    for (int it = 0; it < p->liquid_iterations; it++) {
        int write_pressure = it == p->liquid_iterations - 1;
        /* S/liquid_lambda.comp:36-105 */
        for (int i = 0; i < n; i++) {
            const float *m = s->mats + 4 * i;
            if (m[2] != 0.0f) { lam[i] = 0.0f; continue; }
            float density = 0.0f, ogx = 0.0f, ogy = 0.0f, others = 0.0f;
            for (int q = nstart[i]; q < nstart[i + 1]; q++) {
                int j = nb[q];
                float dx, dy;
                lq_delta(&k, pos + 2 * i, pos + 2 * j, &dx, &dy);
                float d2 = dx * dx + dy * dy;
                if (d2 >= k.h * k.h) continue;
                density += lq_poly6(&k, d2);
                float gx, gy;
                lq_spiky(&k, dx, dy, sqrtf(d2), (uint32_t)i, (uint32_t)j, &gx, &gy);
                gx /= k.rest; gy /= k.rest;
                ogx += gx; ogy += gy;
                if (s->mats[4 * j + 2] == 0.0f) others += gx * gx + gy * gy;
            }
            density += lq_poly6(&k, 0.0f);
            float wr, wgx, wgy;
            lq_walls(&k, pos[2 * i] * k.vw, pos[2 * i + 1] * k.vh, &wr, &wgx, &wgy);
            density += wr;
            ogx += wgx / k.rest; ogy += wgy / k.rest;
            float c = density / k.rest - 1.0f;
            if (c < 0.0f) c = 0.0f;
            float lambda = -c / (ogx * ogx + ogy * ogy + others + k.relax);
            lam[i] = lambda * m[0];
            if (write_pressure) s->press[i] = c * k.pressure_scale;
        }
        /* S/liquid_apply.comp:38-92 */
        for (int i = 0; i < n; i++) {
            const float *b = s->bodies + 4 * i;
            const float *m = s->mats + 4 * i;
            float own_mass = m[0];
            float own_lambda = lam[i] / own_mass;
            float cx = 0.0f, cy = 0.0f;
            for (int q = nstart[i]; q < nstart[i + 1]; q++) {
                int j = nb[q];
                float dx, dy;
                lq_delta(&k, b, pos + 2 * j, &dx, &dy);
                float d2 = dx * dx + dy * dy;
                if (d2 >= k.h * k.h) continue;
                float coef = own_lambda + lam[j] / own_mass;
                if (coef == 0.0f) continue;
                float gx, gy;
                lq_spiky(&k, dx, dy, sqrtf(d2), (uint32_t)i, (uint32_t)j, &gx, &gy);
                cx += coef * gx; cy += coef * gy;
            }
            if (m[2] == 0.0f) {
                float wr, wgx, wgy;
                lq_walls(&k, b[0] * k.vw, b[1] * k.vh, &wr, &wgx, &wgy);
                cx += own_lambda * wgx; cy += own_lambda * wgy;
            }
            cx /= k.rest; cy /= k.rest;
            float ppx = b[0] * k.vw + cx, ppy = b[1] * k.vh + cy;
            float vpx = b[2] * k.vw + cx / p->dt, vpy = b[3] * k.vh + cy / p->dt;
            if (p->bounded) {
                float r = 0.5f * p->diameter;
                float ix = clampf(ppx, r, k.vw - r), iy = clampf(ppy, r, k.vh - r);
                int lowx = ppx < ix, lowy = ppy < iy, highx = ppx > ix, highy = ppy > iy;
                if ((lowx && vpx < 0.0f) || (highx && vpx > 0.0f)) vpx = 0.0f;
                if ((lowy && vpy < 0.0f) || (highy && vpy > 0.0f)) vpy = 0.0f;
                ppx = ix; ppy = iy;
            }
            float *o = out + 4 * i;
            if (p->bounded) {
                o[0] = clampf(ppx / k.vw, 0.0f, MAXIMUM_UNIT);
                o[1] = clampf(ppy / k.vh, 0.0f, MAXIMUM_UNIT);
            } else {
                o[0] = fractf(ppx / k.vw);
                o[1] = fractf(ppy / k.vh);
            }
            o[2] = vpx / k.vw; o[3] = vpy / k.vh;
            pout[2 * i] = o[0]; pout[2 * i + 1] = o[1];
        }
        memcpy(s->bodies, out, (size_t)n * 16);
        memcpy(pos, pout, (size_t)n * 8);
    }
    /* S/liquid_viscosity.comp:27-55 */
    if (p->viscosity > 0.0f) {
        for (int i = 0; i < n; i++) {
            const float *b = s->bodies + 4 * i;
            float *o = out + 4 * i;
            memcpy(o, b, 16);
            if (s->mats[4 * i + 2] != 0.0f) continue;
            float vx = b[2] * k.vw, vy = b[3] * k.vh;
            float bx = 0.0f, by = 0.0f;
            for (int q = nstart[i]; q < nstart[i + 1]; q++) {
                int j = nb[q];
                if (s->mats[4 * j + 2] != 0.0f) continue;
                float dx, dy;
                lq_delta(&k, b, pos + 2 * j, &dx, &dy);
                float d2 = dx * dx + dy * dy;
                if (d2 >= k.h * k.h) continue;
                float w = lq_poly6(&k, d2);
                bx += (s->bodies[4 * j + 2] * k.vw - vx) * w;
                by += (s->bodies[4 * j + 3] * k.vh - vy) * w;
            }
            float c = p->viscosity / k.rest;
            o[2] = (vx + bx * c) / k.vw;
            o[3] = (vy + by * c) / k.vh;
        }
        memcpy(s->bodies, out, (size_t)n * 16);
    }
    free(pos); free(lam); free(out); free(pout); free(nb); free(nstart);
    free(cell); free(start); free(items);
}

/* ------------------------------------------------------------------------ */
/* Reference for the adjoint: the gravity-only leapfrog in FP64             */
/* (K/GravityPipeline.kt:445-494 with drift-only passes, buildForce :907-973, */
/* applyKick :1213-1261). The same P3M force as the FP32 path (Gaussian-split */
/* kernel, CIC deposit and gather, the same deconvolution, the near field of */
/* S/kick.comp with the same topology rules), evaluated in double with the  */
/* CIC deposit unrounded, so its central finite differences give the FP64   */
/* gradient the HIP adjoint is checked against. No spatial sort: it only    */
/* permutes the bodies.                                                     */

typedef struct {
    int mw, mh, pw, ph, n, rx, ry;
    double aspect, G, total_mass, sigma, md2;
    double *tw, *kx, *ky, *rho, *fx, *fy;      /* complex meshes, FP64 spectra of the kernel */
    int *cell_start, *cell_list;
    double *cell_mass;
} grav64;

static void g64_init(grav64 *g, const orc_params *p, int n) {
    g->pw = p->grid_w; g->ph = p->grid_h;
    g->mw = p->bounded ? 2 * g->pw : g->pw;
    g->mh = p->bounded ? 2 * g->ph : g->ph;
    g->n = n;
    g->aspect = (double)p->width / (double)p->height;
    g->G = p->gravity;
    g->total_mass = n > 1 ? n : 1;
    g->sigma = (double)p->softening_cells / (double)g->ph;
    double cw = g->aspect / g->pw, ch = 1.0 / g->ph;
    double md = p->min_dist_frac * (cw < ch ? cw : ch);
    g->md2 = md * md;
    orc_neighbor_radius(p, &g->rx, &g->ry);
    size_t m = (size_t)g->mw * g->mh;
    int max_axis = g->mw > g->mh ? g->mw : g->mh;
    g->tw = malloc(sizeof(double) * 4 * (size_t)max_axis);
    orc_twiddles_f64(max_axis, g->tw);
    g->kx = calloc(2 * m, sizeof(double)); g->ky = calloc(2 * m, sizeof(double));
    g->rho = malloc(sizeof(double) * 2 * m);
    g->fx = malloc(sizeof(double) * 2 * m); g->fy = malloc(sizeof(double) * 2 * m);
    const double PI = 3.141592653589793238463;
    for (size_t id = 0; id < m; id++) {
        int rx = (int)(id % g->mw), ry = (int)(id / g->mw);
        int ox = rx <= g->mw / 2 ? rx : rx - g->mw;
        int oy = ry <= g->mh / 2 ? ry : ry - g->mh;
        double dx = ox * g->aspect / g->pw, dy = (double)oy / g->ph;
        double r2 = dx * dx + dy * dy;
        if (r2 == 0.0) continue;
        double lw = 1.0 - exp(-r2 / (2.0 * g->sigma * g->sigma));
        double dirx = rx * 2 == g->mw ? 0.0 : dx, diry = ry * 2 == g->mh ? 0.0 : dy;
        double scale = -g->G * lw / (g->total_mass * r2);
        g->kx[2 * id] = scale * dirx / g->aspect;
        g->ky[2 * id] = scale * diry;
    }
    orc_fft2d_f64(g->kx, g->mw, g->mh, g->tw, 0);
    orc_fft2d_f64(g->ky, g->mw, g->mh, g->tw, 0);
    /* S/convolve.comp:24-47 folded into the spectra */
    for (size_t id = 0; id < m; id++) {
        int rx = (int)(id % g->mw), ry = (int)(id / g->mw);
        int ikx = rx <= g->mw / 2 ? rx : rx - g->mw, iky = ry <= g->mh / 2 ? ry : ry - g->mh;
        double ax = PI * ikx / g->mw, ay = PI * iky / g->mh;
        double sx = fabs(ax) < 1e-12 ? 1.0 : sin(ax) / ax, sy = fabs(ay) < 1e-12 ? 1.0 : sin(ay) / ay;
        double aw = sx * sx * sy * sy, p2 = aw * aw;
        double s = 1.0 / (p2 > p->window_floor ? p2 : p->window_floor) / (double)m;
        g->kx[2 * id] *= s; g->kx[2 * id + 1] *= s;
        g->ky[2 * id] *= s; g->ky[2 * id + 1] *= s;
    }
    size_t cells = (size_t)g->pw * g->ph;
    g->cell_start = malloc(sizeof(int) * (cells + 1));
    g->cell_list = malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    g->cell_mass = malloc(sizeof(double) * cells);
}

static void g64_free(grav64 *g) {
    free(g->tw); free(g->kx); free(g->ky); free(g->rho); free(g->fx); free(g->fy);
    free(g->cell_start); free(g->cell_list); free(g->cell_mass);
}

static double g64_image(const orc_params *p, double d) { return p->bounded ? d : d - floor(d + 0.5); }

static void g64_local(const grav64 *g, const orc_params *p, double sx, double sy, double tx,
                      double ty, double mass, double *ax, double *ay) {
    double px = g64_image(p, sx - tx) * g->aspect, py = g64_image(p, sy - ty);
    double r2 = px * px + py * py;
    double lw = exp(-r2 / (2.0 * g->sigma * g->sigma));
    double scale = g->G * mass * lw / (g->total_mass * (r2 > g->md2 ? r2 : g->md2));
    *ax += scale * px / g->aspect;
    *ay += scale * py;
}

/* Accelerations of every body at positions xy (normalised), masses m. */
static void g64_accel(grav64 *g, const orc_params *p, const double *xy, const float *mats,
                      double *acc) {
    size_t m = (size_t)g->mw * g->mh;
    for (size_t i = 0; i < 2 * m; i++) g->rho[i] = 0.0;
    size_t cells = (size_t)g->pw * g->ph;
    int *count = calloc(cells, sizeof(int)), *home = malloc(sizeof(int) * (size_t)(g->n > 0 ? g->n : 1));
    for (size_t c = 0; c < cells; c++) g->cell_mass[c] = 0.0;
    for (int i = 0; i < g->n; i++) {
        double x = xy[2 * i], y = xy[2 * i + 1], mass = mats[4 * i];
        double gx = x * g->pw - 0.5, gy = y * g->ph - 0.5;
        double flx = floor(gx), fly = floor(gy);
        int bx = (int)flx, by = (int)fly;
        double fx = gx - flx, fy = gy - fly;
        double w[4] = {(1 - fx) * (1 - fy), fx * (1 - fy), (1 - fx) * fy, fx * fy};
        for (int k = 0; k < 4; k++) {
            int mx = wrap_axis(bx + (k & 1), g->mw), my = wrap_axis(by + (k >> 1), g->mh);
            g->rho[2 * ((size_t)my * g->mw + mx)] += w[k] * mass;
        }
        int hx = (int)floor(x * g->pw), hy = (int)floor(y * g->ph);
        home[i] = wrap_axis(hx, g->pw) + wrap_axis(hy, g->ph) * g->pw;
        count[home[i]]++;
        g->cell_mass[home[i]] += mass;
    }
    g->cell_start[0] = 0;
    for (size_t c = 0; c < cells; c++) g->cell_start[c + 1] = g->cell_start[c] + count[c];
    int *fill = malloc(sizeof(int) * cells);
    for (size_t c = 0; c < cells; c++) fill[c] = g->cell_start[c];
    /* descending body index within a cell, the canonical list order */
    for (int i = g->n - 1; i >= 0; i--) g->cell_list[fill[home[i]]++] = i;
    free(fill);
    orc_fft2d_f64(g->rho, g->mw, g->mh, g->tw, 0);
    for (size_t id = 0; id < m; id++) {
        double rr = g->rho[2 * id], ri = g->rho[2 * id + 1];
        const double *a = g->kx + 2 * id, *b = g->ky + 2 * id;
        g->fx[2 * id] = rr * a[0] - ri * a[1];
        g->fx[2 * id + 1] = rr * a[1] + ri * a[0];
        g->fy[2 * id] = rr * b[0] - ri * b[1];
        g->fy[2 * id + 1] = rr * b[1] + ri * b[0];
    }
    orc_fft2d_f64(g->fx, g->mw, g->mh, g->tw, 1);
    orc_fft2d_f64(g->fy, g->mw, g->mh, g->tw, 1);
    for (int i = 0; i < g->n; i++) {
        double x = xy[2 * i], y = xy[2 * i + 1];
        double gx = x * g->pw - 0.5, gy = y * g->ph - 0.5;
        double flx = floor(gx), fly = floor(gy);
        int bx = (int)flx, by = (int)fly;
        double fx = gx - flx, fy = gy - fly;
        double w[4] = {(1 - fx) * (1 - fy), fx * (1 - fy), (1 - fx) * fy, fx * fy};
        double ax = 0.0, ay = 0.0;
        for (int k = 0; k < 4; k++) {
            int mx = wrap_axis(bx + (k & 1), g->mw), my = wrap_axis(by + (k >> 1), g->mh);
            size_t idx = (size_t)my * g->mw + mx;
            ax += w[k] * g->fx[2 * idx];
            ay += w[k] * g->fy[2 * idx];
        }
        ax += p->field_x / p->width;
        ay += p->field_y / p->height;
        int hx = (int)floor(x * g->pw), hy = (int)floor(y * g->ph);
        for (int dy = -g->ry; dy <= g->ry; dy++) {
            for (int dx = -g->rx; dx <= g->rx; dx++) {
                int cx = hx + dx, cy = hy + dy;
                if (p->bounded && (cx < 0 || cy < 0 || cx >= g->pw || cy >= g->ph)) continue;
                int wx = wrap_axis(cx, g->pw), wy = wrap_axis(cy, g->ph);
                int index = wx + wy * g->pw;
                int cnt = count[index];
                int is_home = dx == 0 && dy == 0;
                if (!is_home && cnt > p->exact_limit) {
                    g64_local(g, p, (wx + 0.5) / g->pw, (wy + 0.5) / g->ph, x, y,
                              g->cell_mass[index], &ax, &ay);
                } else {
                    int limit = is_home ? p->home_limit : p->exact_limit;
                    int visited = 0;
                    for (int t = g->cell_start[index]; t < g->cell_start[index + 1] && visited < limit; t++) {
                        int other = g->cell_list[t];
                        if (other != i)
                            g64_local(g, p, xy[2 * other], xy[2 * other + 1], x, y, mats[4 * other], &ax, &ay);
                        visited++;
                    }
                }
            }
        }
        acc[2 * i] = ax;
        acc[2 * i + 1] = ay;
    }
    free(count); free(home);
}

/* `steps` gravity-only leapfrog steps from bodies (normalised x, y, vx, vy),
 * initial velocities offset by dv (nullable, normalised), in FP64. Writes
 * the final positions (nullable) and returns the loss
 * mean |minimum image((x_T - target) * viewport)|^2 in px^2. */
double orc_grav64_loss(const orc_params *p, int n, const float *bodies, const float *mats,
                       int steps, const float *targets, const double *dv, double *final_xy) {
    grav64 g;
    g64_init(&g, p, n);
    size_t nn = (size_t)(n > 0 ? n : 1);
    double *xy = malloc(16 * nn), *v = malloc(16 * nn), *a = malloc(16 * nn);
    for (int i = 0; i < n; i++) {
        xy[2 * i] = bodies[4 * i]; xy[2 * i + 1] = bodies[4 * i + 1];
        v[2 * i] = bodies[4 * i + 2] + (dv ? dv[2 * i] : 0.0);
        v[2 * i + 1] = bodies[4 * i + 3] + (dv ? dv[2 * i + 1] : 0.0);
    }
    double dt = p->dt, half = 0.5 * p->dt;
    g64_accel(&g, p, xy, mats, a);
    for (int s = 0; s < steps; s++) {
        for (int i = 0; i < 2 * n; i++) v[i] += a[i] * half;
        for (int i = 0; i < 2 * n; i++) {
            double x = xy[i] + v[i] * dt;
            if (p->bounded) x = x < 0.0 ? 0.0 : (x > 1.0 ? 1.0 : x);
            else x -= floor(x);
            xy[i] = x;
        }
        g64_accel(&g, p, xy, mats, a);
        for (int i = 0; i < 2 * n; i++) v[i] += a[i] * half;
    }
    double loss = 0.0;
    for (int i = 0; i < n; i++) {
        double dx = (xy[2 * i] - targets[2 * i]) * p->width;
        double dy = (xy[2 * i + 1] - targets[2 * i + 1]) * p->height;
        if (!p->bounded) {
            dx -= floor(dx / p->width + 0.5) * p->width;
            dy -= floor(dy / p->height + 0.5) * p->height;
        }
        loss += dx * dx + dy * dy;
    }
    if (final_xy) memcpy(final_xy, xy, 16 * (size_t)n);
    free(xy); free(v); free(a);
    g64_free(&g);
    return n > 0 ? loss / n : 0.0;
}

/* ==== glued blocks: the rigid projection ================================== */

#define RG_TILE 256
#define RG_MAXIMUM_GROUPS 256
#define RG_PARTNERS (RG_MAXIMUM_GROUPS + 5)
#define RG_KEYS (RG_MAXIMUM_GROUPS * RG_PARTNERS)

/* S/rigid_layout.glsl:4-15 */
typedef struct rg_group {
    float frame[4], motion[4], scalars[4];
    int32_t tiles[4];
    float shape[4];
    int32_t grid[4];
    float bounds[4];
    int32_t guard[4], support[4];
    float pose[4];
} rg_group;

/* Kotlin's Float.compareTo (-0 below +0) and data-class float equality */
static int rg_before(float a, float b) {
    if (a < b) return 1;
    return a == b && signbit(a) && !signbit(b);
}
static int rg_same(float a, float b) { return a == b && !signbit(a) == !signbit(b); }

/* K/GlueShape.kt:158-190 */
int orc_rigid_hull(const float *xs, const float *ys, int n, int *corners) {
    if (n <= 0) return 0;
    float minY = ys[0], maxY = ys[0];
    for (int i = 1; i < n; i++) {
        if (ys[i] < minY) minY = ys[i];
        if (ys[i] > maxY) maxY = ys[i];
    }
    const int rows = (int)((maxY - minY) / 0.01f) + 1;
    int *lowest = malloc(sizeof(int) * (size_t)rows), *highest = malloc(sizeof(int) * (size_t)rows);
    for (int r = 0; r < rows; r++) lowest[r] = highest[r] = -1;
    for (int i = 0; i < n; i++) {
        const int r = (int)((ys[i] - minY) / 0.01f);
        if (lowest[r] < 0 || xs[i] < xs[lowest[r]]) lowest[r] = i;
        if (highest[r] < 0 || xs[i] > xs[highest[r]]) highest[r] = i;
    }
    int *pts = malloc(sizeof(int) * 2 * (size_t)rows);
    int m = 0;
    for (int r = 0; r < rows; r++) {
        if (lowest[r] < 0) continue;
        const int ends[2] = {lowest[r], highest[r]};
        for (int e = 0; e < 2; e++) {
            int seen = 0;
            for (int k = 0; k < m && !seen; k++)
                seen = rg_same(xs[pts[k]], xs[ends[e]]) && rg_same(ys[pts[k]], ys[ends[e]]);
            if (!seen) pts[m++] = ends[e];
        }
    }
    /* stable insertion sort by (x, y) */
    for (int i = 1; i < m; i++) {
        const int v = pts[i];
        int j = i - 1;
        while (j >= 0 && (rg_before(xs[v], xs[pts[j]]) ||
                          (rg_same(xs[v], xs[pts[j]]) && rg_before(ys[v], ys[pts[j]])))) {
            pts[j + 1] = pts[j];
            j--;
        }
        pts[j + 1] = v;
    }
    int count = 0;
    if (m < 3) {
        for (int k = 0; k < m; k++) corners[count++] = pts[k];
    } else {
        /* Andrew's monotone chain, lower then upper, collinear points dropped */
        int *chain = malloc(sizeof(int) * 2 * (size_t)(m + 1));
        int len = 0;
        for (int pass = 0; pass < 2; pass++) {
            const int start = len;
            for (int k = 0; k < m; k++) {
                const int q = pass == 0 ? pts[k] : pts[m - 1 - k];
                while (len >= start + 2) {
                    const int o = chain[len - 2], a = chain[len - 1];
                    const double turn = (double)(xs[a] - xs[o]) * (double)(ys[q] - ys[o]) -
                                        (double)(ys[a] - ys[o]) * (double)(xs[q] - xs[o]);
                    if (turn > 0.0) break;
                    len--;
                }
                chain[len++] = q;
            }
            len--;
        }
        for (int k = 0; k < len; k++) corners[count++] = chain[k];
        free(chain);
    }
    free(pts);
    free(lowest);
    free(highest);
    return count;
}

/* K/RigidBlocks.kt:718-721, 751-777, 838-840 */
void orc_rigid_prepare(const float *offsets_xy, int n, float *rest_xy, float *centroid_xy,
                       double *inertia, float *bounds) {
    double sx = 0.0, sy = 0.0;
    for (int i = 0; i < n; i++) {
        sx += (double)offsets_xy[2 * i];
        sy += (double)offsets_xy[2 * i + 1];
    }
    const float cx = (float)(sx / n), cy = (float)(sy / n);
    double sum = 0.0;
    float lx = offsets_xy[0], hx = offsets_xy[0], ly = offsets_xy[1], hy = offsets_xy[1];
    for (int i = 0; i < n; i++) {
        const float x = offsets_xy[2 * i] - cx, y = offsets_xy[2 * i + 1] - cy;
        rest_xy[2 * i] = x;
        rest_xy[2 * i + 1] = y;
        const float term = x * x + y * y;
        sum += (double)term;
        if (offsets_xy[2 * i] < lx) lx = offsets_xy[2 * i];
        if (offsets_xy[2 * i] > hx) hx = offsets_xy[2 * i];
        if (offsets_xy[2 * i + 1] < ly) ly = offsets_xy[2 * i + 1];
        if (offsets_xy[2 * i + 1] > hy) hy = offsets_xy[2 * i + 1];
    }
    centroid_xy[0] = cx;
    centroid_xy[1] = cy;
    *inertia = sum;
    bounds[0] = lx - cx;
    bounds[1] = hx - cx;
    bounds[2] = ly - cy;
    bounds[3] = hy - cy;
}

/* S/rigid_reduce.comp:60-69: the workgroup's tree, stride 128 down to 1 */
static void rg_tree(float (*a)[4], float (*b)[4]) {
    for (int stride = RG_TILE / 2; stride > 0; stride >>= 1)
        for (int l = 0; l < stride; l++)
            for (int c = 0; c < 4; c++) {
                a[l][c] = a[l][c] + a[l + stride][c];
                b[l][c] = b[l][c] + b[l + stride][c];
            }
}

/* S/rigid_reduce.comp:42-58 */
static void rg_stop_against(float *v, float px, float py, float friction, float substep_dt) {
    const float distance = sqrtf(px * px + py * py);
    if (distance == 0.0f) return;
    const float dx = px / distance, dy = py / distance;
    const float against = v[0] * dx + v[1] * dy;
    if (against >= 0.0f) return;
    const float stop = fminf(-against, distance / substep_dt);
    v[0] = v[0] + dx * stop;
    v[1] = v[1] + dy * stop;
    const float along = v[0] * dx + v[1] * dy;
    const float sx = v[0] - along * dx, sy = v[1] - along * dy;
    const float sliding = sqrtf(sx * sx + sy * sy);
    if (sliding > 0.0f) {
        const float f = fminf(sliding, friction * stop) / sliding;
        v[0] = v[0] - sx * f;
        v[1] = v[1] - sy * f;
    }
}

/* S/rigid_reduce.comp:78-100 */
static void rg_settle(int32_t *memory, int group, int group_count) {
    int32_t *impacts = memory, *next = memory + RG_KEYS, *states = memory + 2 * RG_KEYS;
    int32_t *idle = memory + 3 * RG_KEYS, *pairs = memory + 4 * RG_KEYS;
    const int partners = group_count - group;
    for (int index = 0; index < partners + 5; index++) {
        const int partner = index < partners ? group + index : RG_MAXIMUM_GROUPS + index - partners;
        const int key = group * RG_PARTNERS + partner;
        if (states[key] == 2) {
            int speed = impacts[key] & 0xffffff;
            if (next[key] > speed) speed = next[key];
            const int age = (impacts[key] >> 24) + 1;
            impacts[key] = speed == 0 || age > 32 ? 0 : speed | (age << 24);
            idle[key] = 0;
        } else if (states[key] == 1) {
            impacts[key] = 0;
            idle[key] = 0;
        } else if (impacts[key] != 0 && ++idle[key] >= 4) {
            impacts[key] = 0;
            idle[key] = 0;
        }
        next[key] = 0;
        states[key] = 0;
        pairs[key] = 0;
    }
}

void orc_rigid_project(const orc_rigid_params *p, void *groups_words, int32_t *memory, const float *shapes,
                       const int32_t *members, const float *rest, const int32_t *tile_groups,
                       float *bodies) {
    rg_group *groups = (rg_group *)groups_words;
    const float W = p->width, H = p->height;
    float (*partials)[2][4] = calloc((size_t)(p->tile_count > 0 ? p->tile_count : 1), sizeof *partials);
    float (*sa)[4] = malloc(sizeof(float[4]) * RG_TILE), (*sb)[4] = malloc(sizeof(float[4]) * RG_TILE);
    /* S/rigid_reduce.comp:203-229: the tile sums */
    for (int tile = 0; tile < p->tile_count; tile++) {
        const rg_group *r = &groups[tile_groups[tile]];
        if (tile < r->tiles[0] || tile >= r->tiles[0] + r->tiles[1]) continue;
        for (int lane = 0; lane < RG_TILE; lane++) {
            const int member = tile * RG_TILE + lane;
            const int b = members[member];
            for (int c = 0; c < 4; c++) sa[lane][c] = sb[lane][c] = 0.0f;
            if (b < 0) continue;
            const float *state = bodies + 4 * (size_t)b;
            float px = state[0] * W - r->frame[0], py = state[1] * H - r->frame[1];
            if (!p->bounded) {
                px -= floorf(px / W + 0.5f) * W;
                py -= floorf(py / H + 0.5f) * H;
            }
            const float vx = state[2] * W, vy = state[3] * H;
            const float ox = rest[2 * member], oy = rest[2 * member + 1];
            const float rx = r->motion[2], ry = r->motion[3];
            const float ax = rx * ox - ry * oy, ay = ry * ox + rx * oy;
            sa[lane][0] = px; sa[lane][1] = py; sa[lane][2] = vx; sa[lane][3] = vy;
            sb[lane][0] = ox * py - oy * px;
            sb[lane][1] = ox * px + oy * py;
            sb[lane][2] = ax * vy - ay * vx;
        }
        rg_tree(sa, sb);
        memcpy(partials[tile][0], sa[0], sizeof(float[4]));
        memcpy(partials[tile][1], sb[0], sizeof(float[4]));
    }
    /* S/rigid_reduce.comp:102-201: the fit of every block */
    for (int g = 0; g < p->group_count; g++) {
        rg_group *r = &groups[g];
        if (r->tiles[1] == 0) continue;
        for (int lane = 0; lane < RG_TILE; lane++) {
            for (int c = 0; c < 4; c++) sa[lane][c] = sb[lane][c] = 0.0f;
            for (int tile = r->tiles[0] + lane; tile < r->tiles[0] + r->tiles[1]; tile += RG_TILE)
                for (int c = 0; c < 4; c++) {
                    sa[lane][c] = sa[lane][c] + partials[tile][0][c];
                    sb[lane][c] = sb[lane][c] + partials[tile][1][c];
                }
        }
        rg_tree(sa, sb);
        if (p->contacts) rg_settle(memory, g, p->group_count);
        const float mass = r->scalars[2];
        if (mass <= 0.0f) continue;
        float cx = sa[0][0] / mass, cy = sa[0][1] / mass;
        float v[2] = {sa[0][2] / mass, sa[0][3] / mass};
        const float fx = sb[0][1], fy = sb[0][0];
        const float fit = sqrtf(fx * fx + fy * fy);
        const float rx = fit > 0.0f ? fx / fit : r->motion[2];
        const float ry = fit > 0.0f ? fy / fit : r->motion[3];
        const float spin = r->scalars[1] > 0.0f ? sb[0][2] / r->scalars[1] : 0.0f;
        cx += r->frame[0];
        cy += r->frame[1];
        cx += (float)(r->guard[0] - r->guard[2]) / 65536.0f;
        cy += (float)(r->guard[1] - r->guard[3]) / 65536.0f;
        const float sx = (float)(r->support[0] - r->support[2]) / 65536.0f;
        const float sy = (float)(r->support[1] - r->support[3]) / 65536.0f;
        for (int c = 0; c < 4; c++) r->guard[c] = r->support[c] = 0;
        cx += sx;
        cy += sy;
        if (p->contacts) rg_stop_against(v, sx, sy, 0.0f, p->substep_dt);
        if (!p->bounded) {
            cx -= floorf(cx / W) * W;
            cy -= floorf(cy / H) * H;
        } else {
            float lx = 1e30f, ly = 1e30f, hx = -1e30f, hy = -1e30f;
            for (int k = 0; k < r->tiles[3]; k++) {
                const float ox = shapes[r->tiles[2] + 2 * k], oy = shapes[r->tiles[2] + 2 * k + 1];
                const float ax = rx * ox - ry * oy, ay = ry * ox + rx * oy;
                lx = fminf(lx, ax); ly = fminf(ly, ay);
                hx = fmaxf(hx, ax); hy = fmaxf(hy, ay);
            }
            const float rad = p->body_radius;
            float shx = fmaxf(rad - (cx + lx), 0.0f) - fmaxf(cx + hx + rad - W, 0.0f);
            float shy = fmaxf(rad - (cy + ly), 0.0f) - fmaxf(cy + hy + rad - H, 0.0f);
            shx *= hx - lx + 2.0f * rad < W ? 1.0f : 0.0f;
            shy *= hy - ly + 2.0f * rad < H ? 1.0f : 0.0f;
            cx += shx;
            cy += shy;
            if (p->contacts) rg_stop_against(v, shx, shy, p->block_friction, p->substep_dt);
        }
        r->frame[0] = r->frame[2] = cx;
        r->frame[1] = r->frame[3] = cy;
        r->motion[0] = v[0]; r->motion[1] = v[1]; r->motion[2] = rx; r->motion[3] = ry;
        r->scalars[0] = spin;
    }
    /* S/rigid_apply.comp:48-69: every member onto its block's motion */
    for (int tile = 0; tile < p->tile_count; tile++) {
        const rg_group *r = &groups[tile_groups[tile]];
        for (int lane = 0; lane < RG_TILE; lane++) {
            const int member = tile * RG_TILE + lane;
            const int b = members[member];
            if (b < 0) continue;
            const float ox = rest[2 * member], oy = rest[2 * member + 1];
            const float rx = r->motion[2], ry = r->motion[3];
            const float ax = rx * ox - ry * oy, ay = ry * ox + rx * oy;
            float px = (r->frame[2] + ax) / W, py = (r->frame[3] + ay) / H;
            const float vx = r->motion[0] + r->scalars[0] * -ay, vy = r->motion[1] + r->scalars[0] * ax;
            if (p->bounded) {
                px = fminf(fmaxf(px, 0.0f), 0.99999994f);
                py = fminf(fmaxf(py, 0.0f), 0.99999994f);
            } else {
                px = px - floorf(px);
                py = py - floorf(py);
            }
            float *out = bodies + 4 * (size_t)b;
            out[0] = px; out[1] = py; out[2] = vx / W; out[3] = vy / H;
        }
    }
    free(partials);
    free(sa);
    free(sb);
}
