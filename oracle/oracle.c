/*
 * oracle.c -- CPU restatement of the reference Barnes-Hut disk step.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Built with -O2 -ffp-contract=off
 * and without -ffast-math so that every float expression rounds exactly as
 * written (one rounding per operation, no fused multiply-add).
 *
 * Citations are to the reference repository ualiberik/nbody-sim.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_MAX_NODES_MULT 8 /* BHTree.cuh:4 */

/* ======================================================================== */
/* Reference canonical octree                                                */
/* ======================================================================== */

orc_tree *orc_tree_new(int n)
{
    orc_tree *t = (orc_tree *)calloc(1, sizeof(orc_tree));
    if (!t) return NULL;
    t->n_bodies = n;
    t->n_total = n + ORC_MAX_NODES_MULT * n;
    size_t nt = (size_t)t->n_total;
    t->pos_x = (float *)calloc(nt, sizeof(float));
    t->pos_y = (float *)calloc(nt, sizeof(float));
    t->pos_z = (float *)calloc(nt, sizeof(float));
    t->mass = (float *)calloc(nt, sizeof(float));
    t->cell_size = (float *)calloc(nt, sizeof(float));
    t->child = (int *)malloc(nt * 8 * sizeof(int));
    t->count = (int *)calloc(nt, sizeof(int));
    if (!t->pos_x || !t->pos_y || !t->pos_z || !t->mass || !t->cell_size || !t->child ||
        !t->count) {
        orc_tree_free(t);
        return NULL;
    }
    return t;
}

void orc_tree_free(orc_tree *t)
{
    if (!t) return;
    free(t->pos_x);
    free(t->pos_y);
    free(t->pos_z);
    free(t->mass);
    free(t->cell_size);
    free(t->child);
    free(t->count);
    free(t);
}

/* resetOctree (BHTree.cu:42-71, 397-420): zero geometry, child = -1. */
static void orc_reset(orc_tree *t)
{
    size_t nt = (size_t)t->n_total;
    memset(t->pos_x, 0, nt * sizeof(float));
    memset(t->pos_y, 0, nt * sizeof(float));
    memset(t->pos_z, 0, nt * sizeof(float));
    memset(t->mass, 0, nt * sizeof(float));
    memset(t->cell_size, 0, nt * sizeof(float));
    memset(t->count, 0, nt * sizeof(int));
    for (size_t i = 0; i < nt * 8; ++i) t->child[i] = -1;
    t->next_node = t->n_bodies;
    t->lost = 0;
    t->max_depth = 0;
}

/* BHTree.cu:94-133 (min/max reduction) and :440-446 (host padding). */
void orc_ref_bbox(orc_tree *t, const float *x, const float *y, const float *z, int n)
{
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
    float hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int i = 0; i < n; ++i) {
        lo[0] = fminf(lo[0], x[i]);
        lo[1] = fminf(lo[1], y[i]);
        lo[2] = fminf(lo[2], z[i]);
        hi[0] = fmaxf(hi[0], x[i]);
        hi[1] = fmaxf(hi[1], y[i]);
        hi[2] = fmaxf(hi[2], z[i]);
    }
    for (int k = 0; k < 3; ++k) {
        float mid = (lo[k] + hi[k]) * 0.5f;
        float half = (hi[k] - lo[k]) * 0.5f * 1.001f;
        t->bmin[k] = mid - half;
        t->bmax[k] = mid + half;
    }
}

/* BHTree.cu:164-170 */
static int orc_octant(float cx, float cy, float cz, float px, float py, float pz)
{
    return ((px > cx) ? 1 : 0) | ((py > cy) ? 2 : 0) | ((pz > cz) ? 4 : 0);
}

/* BHTree.cu:147-159 (leaf copy), :139-142 (leaf count), :475-491 (root
 * geometry), :175-277 (insertion), run one body at a time. */
int orc_ref_build(orc_tree *t, const float *x, const float *y, const float *z, const float *m,
                  int n)
{
    float bmin[3] = {t->bmin[0], t->bmin[1], t->bmin[2]};
    float bmax[3] = {t->bmax[0], t->bmax[1], t->bmax[2]};
    orc_reset(t);
    t->bmin[0] = bmin[0]; t->bmin[1] = bmin[1]; t->bmin[2] = bmin[2];
    t->bmax[0] = bmax[0]; t->bmax[1] = bmax[1]; t->bmax[2] = bmax[2];
    for (int i = 0; i < n; ++i) {
        t->pos_x[i] = x[i];
        t->pos_y[i] = y[i];
        t->pos_z[i] = z[i];
        t->mass[i] = m[i];
        t->count[i] = 1;
    }
    const float rcx = (bmin[0] + bmax[0]) * 0.5f;
    const float rcy = (bmin[1] + bmax[1]) * 0.5f;
    const float rcz = (bmin[2] + bmax[2]) * 0.5f;
    const float rhs = fmaxf(fmaxf(bmax[0] - bmin[0], bmax[1] - bmin[1]), bmax[2] - bmin[2]) *
                      0.5f * 1.001f;
    const int root = t->n_total - 1;
    t->pos_x[root] = rcx;
    t->pos_y[root] = rcy;
    t->pos_z[root] = rcz;
    t->cell_size[root] = rhs;

    for (int body = 0; body < n; ++body) {
        const float px = x[body], py = y[body], pz = z[body];
        int node = root, depth = 0;
        float ncx = rcx, ncy = rcy, ncz = rcz, nhs = rhs;
        int max_iter = 128, placed = 0;
        while (max_iter-- > 0) {
            int oct = orc_octant(ncx, ncy, ncz, px, py, pz);
            int *slot = &t->child[node * 8 + oct];
            int old = *slot;
            if (old == -1) {
                *slot = body;
                placed = 1;
                break;
            }
            if (old >= n) { /* descend (BHTree.cu:226-232) */
                node = old;
                ncx = t->pos_x[node];
                ncy = t->pos_y[node];
                ncz = t->pos_z[node];
                nhs = t->cell_size[node];
                ++depth;
                continue;
            }
            /* subdivide (BHTree.cu:244-274) */
            int new_node = t->next_node++;
            if (new_node >= t->n_total) break; /* pool overflow: body lost */
            float child_hs = nhs * 0.5f;
            float new_cx = ncx + ((px > ncx) ? child_hs : -child_hs);
            float new_cy = ncy + ((py > ncy) ? child_hs : -child_hs);
            float new_cz = ncz + ((pz > ncz) ? child_hs : -child_hs);
            t->pos_x[new_node] = new_cx;
            t->pos_y[new_node] = new_cy;
            t->pos_z[new_node] = new_cz;
            t->cell_size[new_node] = child_hs;
            int disp = orc_octant(new_cx, new_cy, new_cz, t->pos_x[old], t->pos_y[old],
                                  t->pos_z[old]);
            t->child[new_node * 8 + disp] = old;
            *slot = new_node;
            node = new_node;
            ncx = new_cx;
            ncy = new_cy;
            ncz = new_cz;
            nhs = child_hs;
            ++depth;
        }
        if (!placed) ++t->lost;
        if (depth > t->max_depth) t->max_depth = depth;
    }
    return t->lost;
}

/* BHTree.cu:292-316 applied bottom-up to completion.  Internal nodes are
 * allocated after their parent, so descending index order visits children
 * first; the root (highest index) is last. */
static void orc_summarize_node(orc_tree *t, int node)
{
    float m = 0.f, cx = 0.f, cy = 0.f, cz = 0.f;
    int cnt = 0;
    for (int k = 0; k < 8; ++k) {
        int c = t->child[node * 8 + k];
        if (c < 0) continue;
        float cm = t->mass[c];
        m += cm;
        cx += t->pos_x[c] * cm;
        cy += t->pos_y[c] * cm;
        cz += t->pos_z[c] * cm;
        cnt += t->count[c];
    }
    if (m > 0.f) {
        t->pos_x[node] = cx / m;
        t->pos_y[node] = cy / m;
        t->pos_z[node] = cz / m;
        t->mass[node] = m;
        t->count[node] = cnt;
    }
}

void orc_ref_summarize(orc_tree *t)
{
    int top = t->next_node < t->n_total ? t->next_node : t->n_total - 1;
    for (int node = top - 1; node >= t->n_bodies; --node) orc_summarize_node(t, node);
    orc_summarize_node(t, t->n_total - 1);
}

/* BHTree.cu:322-352 */
int orc_ref_sort(const orc_tree *t, int32_t *sort_idx)
{
    int *stack = (int *)malloc(2048 * sizeof(int));
    int top = 0, out = 0, dropped = 0;
    stack[top++] = t->n_total - 1;
    while (top > 0) {
        int node = stack[--top];
        if (node < 0) continue;
        if (node < t->n_bodies) {
            sort_idx[out++] = node;
            continue;
        }
        for (int k = 7; k >= 0; --k) {
            int c = t->child[node * 8 + k];
            if (c >= 0) {
                if (top < 2047)
                    stack[top++] = c;
                else
                    ++dropped;
            }
        }
    }
    free(stack);
    return dropped ? -dropped : out;
}

/* BHTree.cu:539-597 for one body. */
static int orc_ref_force_one(const orc_tree *t, int body, float bx, float by, float bz,
                             float theta, float softening, float star_mass, float out[6],
                             int32_t *nacc, int32_t *nopen)
{
    float fx = 0.f, fy = 0.f, fz = 0.f;
    float sx = 0.f, sy = 0.f, sz = 0.f;
    {
        float dx = -bx, dy = -by, dz = -bz;
        float r2 = dx * dx + dy * dy + dz * dz + softening * softening;
        float r = sqrtf(r2);
        float f = star_mass / (r2 * r);
        fx += f * dx;
        fy += f * dy;
        fz += f * dz;
    }
    int stack[256];
    int top = 0, overflow = 0, acc = 0, opn = 0;
    const int n = t->n_bodies;
    stack[top++] = t->n_total - 1;
    while (top > 0) {
        int node = stack[--top];
        if (node < 0) continue;
        float dx = t->pos_x[node] - bx;
        float dy = t->pos_y[node] - by;
        float dz = t->pos_z[node] - bz;
        float r_geom2 = dx * dx + dy * dy + dz * dz;
        float r_geom = sqrtf(r_geom2);
        float r2 = r_geom2 + softening * softening;
        float r = sqrtf(r2);
        int is_leaf = node < n;
        int accept = is_leaf || (t->cell_size[node] / r_geom < theta);
        if (accept) {
            if (node == body) continue;
            float mm = t->mass[node];
            float f = mm / (r2 * r);
            fx += f * dx;
            fy += f * dy;
            fz += f * dz;
            sx += f * dx;
            sy += f * dy;
            sz += f * dz;
            ++acc;
        } else {
            ++opn;
            for (int k = 0; k < 8; ++k) {
                int c = t->child[node * 8 + k];
                if (c >= 0) {
                    if (top < 255)
                        stack[top++] = c;
                    else
                        ++overflow;
                }
            }
        }
    }
    out[0] = fx; out[1] = fy; out[2] = fz;
    out[3] = sx; out[4] = sy; out[5] = sz;
    if (nacc) *nacc = acc;
    if (nopen) *nopen = opn;
    return overflow;
}

int orc_ref_force(const orc_tree *t, const float *x, const float *y, const float *z, int n,
                  float theta, float softening, float star_mass, float *ax, float *ay,
                  float *az, float *self_ax, float *self_ay, float *self_az, int32_t *n_acc,
                  int32_t *n_open, int nthreads)
{
    int overflow = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : overflow) num_threads(nthreads)
#endif
    for (int i = 0; i < n; ++i) {
        float out[6];
        int32_t a = 0, o = 0;
        overflow += orc_ref_force_one(t, i, x[i], y[i], z[i], theta, softening, star_mass, out,
                                      &a, &o);
        if (ax) { ax[i] = out[0]; ay[i] = out[1]; az[i] = out[2]; }
        if (self_ax) { self_ax[i] = out[3]; self_ay[i] = out[4]; self_az[i] = out[5]; }
        if (n_acc) n_acc[i] = a;
        if (n_open) n_open[i] = o;
    }
    return overflow;
}

/* BHTree.cu:628-672 (pass 1) and :676-686 (pass 2). */
long orc_ref_collide(const orc_tree *t, const float *x, const float *y, const float *z,
                     float *vx, float *vy, float *vz, int n, float r_coll, float e, int exact,
                     int nthreads)
{
    float *dv = (float *)calloc((size_t)n * 3, sizeof(float));
    long pairs = 0;
    const float rc2 = r_coll * r_coll;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : pairs) num_threads(nthreads)
#endif
    for (int i = 0; i < n; ++i) {
        float ix = x[i], iy = y[i], iz = z[i];
        float vix = vx[i], viy = vy[i], viz = vz[i];
        float dvx = 0.f, dvy = 0.f, dvz = 0.f;
        int stack[256];
        int top = 0;
        stack[top++] = t->n_total - 1;
        while (top > 0) {
            int node = stack[--top];
            if (node < 0) continue;
            float dx = t->pos_x[node] - ix, dy = t->pos_y[node] - iy, dz = t->pos_z[node] - iz;
            float dist2 = dx * dx + dy * dy + dz * dz;
            if (node < n) {
                int j = node;
                if (j == i) continue;
                if (dist2 < rc2) {
                    ++pairs;
                    float dist = sqrtf(dist2) + 1e-10f;
                    float nx = dx / dist, ny = dy / dist, nz = dz / dist;
                    float dvn = (vx[j] - vix) * nx + (vy[j] - viy) * ny + (vz[j] - viz) * nz;
                    if (dvn < 0.f) {
                        float J = -(1.f + e) * dvn * 0.5f;
                        dvx += -J * nx;
                        dvy += -J * ny;
                        dvz += -J * nz;
                    }
                }
            } else {
                float h = t->cell_size[node];
                int open;
                if (exact) /* every body of the cell lies within 2*sqrt(3)*h of its COM */
                    open = sqrtf(dist2) - 3.5f * h < r_coll * 1.01f;
                else
                    open = sqrtf(dist2) - h * 1.732f < r_coll;
                if (open)
                    for (int k = 0; k < 8; ++k) {
                        int c = t->child[node * 8 + k];
                        if (c >= 0 && top < 255) stack[top++] = c;
                    }
            }
        }
        dv[3 * i] = dvx;
        dv[3 * i + 1] = dvy;
        dv[3 * i + 2] = dvz;
    }
    for (int i = 0; i < n; ++i) {
        vx[i] += dv[3 * i];
        vy[i] += dv[3 * i + 1];
        vz[i] += dv[3 * i + 2];
    }
    free(dv);
    return pairs;
}

/* ======================================================================== */
/* Integrator and perturber                                                  */
/* ======================================================================== */

void orc_half_kick(float *vx, float *vy, float *vz, const float *ax, const float *ay,
                   const float *az, float dt_half, int n)
{
    for (int i = 0; i < n; ++i) {
        vx[i] += ax[i] * dt_half;
        vy[i] += ay[i] * dt_half;
        vz[i] += az[i] * dt_half;
    }
}

void orc_drift(float *x, float *y, float *z, const float *vx, const float *vy, const float *vz,
               float dt, int n)
{
    for (int i = 0; i < n; ++i) {
        x[i] += vx[i] * dt;
        y[i] += vy[i] * dt;
        z[i] += vz[i] * dt;
    }
}

void orc_gas_drag(const float *x, const float *y_unused, const float *z, float *vx, float *vy,
                  float *vz, float drag_rate, float dt, float star_mass, int n)
{
    (void)y_unused;
    if (drag_rate <= 0.0f) return; /* Integrator.cu:102 */
    const float decay = expf(-drag_rate * dt);
    for (int i = 0; i < n; ++i) {
        float xi = x[i], zi = z[i];
        float r_xz2 = xi * xi + zi * zi;
        if (r_xz2 < 1e-12f) continue;
        float r_xz = sqrtf(r_xz2);
        float inv_r = 1.0f / r_xz;
        float vxi = vx[i], vyi = vy[i], vzi = vz[i];
        float v_r = (vxi * xi + vzi * zi) * inv_r;
        float v_phi = (vxi * (-zi) + vzi * xi) * inv_r;
        float v_kep_mag = sqrtf(star_mass * inv_r);
        float v_kep = (v_phi >= 0.0f) ? v_kep_mag : -v_kep_mag;
        float v_r_new = v_r * decay;
        float v_phi_new = v_kep + (v_phi - v_kep) * decay;
        float v_y_new = vyi * decay;
        vx[i] = v_r_new * (xi * inv_r) + v_phi_new * (-zi * inv_r);
        vz[i] = v_r_new * (zi * inv_r) + v_phi_new * (xi * inv_r);
        vy[i] = v_y_new;
    }
}

void orc_perturber_force(float *ax, float *ay, float *az, const float *x, const float *y,
                         const float *z, float px, float py, float pz, float pmass,
                         float softening, int n)
{
    for (int i = 0; i < n; ++i) {
        float dx = px - x[i];
        float dy = py - y[i];
        float dz = pz - z[i];
        float r2 = dx * dx + dy * dy + dz * dz + softening * softening;
        float r = sqrtf(r2);
        float f = pmass / (r2 * r);
        ax[i] += f * dx;
        ay[i] += f * dy;
        az[i] += f * dz;
    }
}

void orc_make_perturber(orc_perturber *p, int enabled, float mass_mjup, float radius_au,
                        float phase0, float star_mass)
{
    memset(p, 0, sizeof(*p));
    p->enabled = enabled;
    if (!enabled) return;
    p->r = radius_au;
    p->mass_msun = mass_mjup * 9.5479e-4f; /* Perturber.h:25 */
    p->phase0 = phase0;
    p->omega = sqrtf(star_mass / (p->r * p->r * p->r));
    p->x = p->r * cosf(p->phase0);
    p->z = p->r * sinf(p->phase0);
    p->y = 0.f;
}

void orc_update_perturber(orc_perturber *p, float sim_time)
{
    if (!p->enabled) return;
    float phi = p->phase0 + p->omega * sim_time;
    p->x = p->r * cosf(phi);
    p->z = p->r * sinf(phi);
    p->y = 0.f;
}

float orc_particle_physical_radius(float particle_mass_msun)
{
    const float PI = 3.14159265358979f, AU_M = 1.496e11f, MSUN_KG = 1.989e30f;
    float mass_kg = particle_mass_msun * MSUN_KG;
    float density = 3000.0f;
    float vol = mass_kg / density;
    float r_m = cbrtf(3.0f * vol / (4.0f * PI));
    return r_m / AU_M;
}

/* ======================================================================== */
/* FoF                                                                       */
/* ======================================================================== */

/* Aggregate.cpp:9-15 */
static int uf_find(int *parent, int i)
{
    while (parent[i] != i) {
        parent[i] = parent[parent[i]];
        i = parent[i];
    }
    return i;
}

/* Aggregate.cpp:17-24 */
static void uf_unite(int *parent, int *rank, int a, int b)
{
    a = uf_find(parent, a);
    b = uf_find(parent, b);
    if (a == b) return;
    if (rank[a] < rank[b]) { int s = a; a = b; b = s; }
    parent[b] = a;
    if (rank[a] == rank[b]) ++rank[a];
}

/* Replace roots by the smallest member index; count groups of size >= 2. */
static int uf_canonical(int *parent, int n, int32_t *labels)
{
    int *minidx = (int *)malloc((size_t)n * sizeof(int));
    int *size = (int *)calloc((size_t)n, sizeof(int));
    for (int i = 0; i < n; ++i) minidx[i] = n;
    for (int i = 0; i < n; ++i) {
        int r = uf_find(parent, i);
        if (i < minidx[r]) minidx[r] = i;
        size[r]++;
    }
    int groups = 0;
    for (int i = 0; i < n; ++i) {
        int r = uf_find(parent, i);
        labels[i] = minidx[r];
        if (r == i && size[r] >= 2) ++groups;
    }
    free(minidx);
    free(size);
    return groups;
}

int orc_fof_n2(const float *x, const float *y, const float *z, int n, float ll,
               int32_t *labels)
{
    float ll2 = ll * ll;
    int *parent = (int *)malloc((size_t)n * sizeof(int));
    int *rank = (int *)calloc((size_t)n, sizeof(int));
    for (int i = 0; i < n; ++i) parent[i] = i;
    for (int i = 0; i < n - 1; ++i)
        for (int j = i + 1; j < n; ++j) {
            float dx = x[i] - x[j], dy = y[i] - y[j], dz = z[i] - z[j];
            if (dx * dx + dy * dy + dz * dz < ll2) uf_unite(parent, rank, i, j);
        }
    int g = uf_canonical(parent, n, labels);
    free(parent);
    free(rank);
    return g;
}

typedef struct { uint64_t cell; int idx; } orc_cellrec;

static int cmp_cellrec(const void *a, const void *b)
{
    const orc_cellrec *p = (const orc_cellrec *)a, *q = (const orc_cellrec *)b;
    if (p->cell != q->cell) return p->cell < q->cell ? -1 : 1;
    return p->idx - q->idx;
}

static uint64_t pack_cell(long cx, long cy, long cz)
{
    return ((uint64_t)(cx & 0x1FFFFF) << 42) | ((uint64_t)(cy & 0x1FFFFF) << 21) |
           (uint64_t)(cz & 0x1FFFFF);
}

int orc_fof_grid(const float *x, const float *y, const float *z, int n, float ll,
                 int32_t *labels)
{
    if (n <= 0) return 0;
    float lo[3] = {x[0], y[0], z[0]};
    for (int i = 1; i < n; ++i) {
        lo[0] = fminf(lo[0], x[i]);
        lo[1] = fminf(lo[1], y[i]);
        lo[2] = fminf(lo[2], z[i]);
    }
    const double cs = (double)ll * 1.01 + 1e-30;
    const float ll2 = ll * ll;
    orc_cellrec *rec = (orc_cellrec *)malloc((size_t)n * sizeof(orc_cellrec));
    long *ci = (long *)malloc((size_t)n * 3 * sizeof(long));
    for (int i = 0; i < n; ++i) {
        long a = (long)floor(((double)x[i] - lo[0]) / cs);
        long b = (long)floor(((double)y[i] - lo[1]) / cs);
        long c = (long)floor(((double)z[i] - lo[2]) / cs);
        ci[3 * i] = a; ci[3 * i + 1] = b; ci[3 * i + 2] = c;
        rec[i].cell = pack_cell(a, b, c);
        rec[i].idx = i;
    }
    qsort(rec, (size_t)n, sizeof(orc_cellrec), cmp_cellrec);
    int *parent = (int *)malloc((size_t)n * sizeof(int));
    int *rank = (int *)calloc((size_t)n, sizeof(int));
    for (int i = 0; i < n; ++i) parent[i] = i;
    for (int i = 0; i < n; ++i) {
        for (int da = -1; da <= 1; ++da)
            for (int db = -1; db <= 1; ++db)
                for (int dc = -1; dc <= 1; ++dc) {
                    long a = ci[3 * i] + da, b = ci[3 * i + 1] + db, c = ci[3 * i + 2] + dc;
                    if (a < 0 || b < 0 || c < 0) continue;
                    uint64_t key = pack_cell(a, b, c);
                    int lo_i = 0, hi_i = n;
                    while (lo_i < hi_i) {
                        int mid = (lo_i + hi_i) / 2;
                        if (rec[mid].cell < key) lo_i = mid + 1; else hi_i = mid;
                    }
                    for (int k = lo_i; k < n && rec[k].cell == key; ++k) {
                        int j = rec[k].idx;
                        if (j <= i) continue;
                        float dx = x[i] - x[j], dy = y[i] - y[j], dz = z[i] - z[j];
                        if (dx * dx + dy * dy + dz * dz < ll2) uf_unite(parent, rank, i, j);
                    }
                }
    }
    int g = uf_canonical(parent, n, labels);
    free(rec);
    free(ci);
    free(parent);
    free(rank);
    return g;
}

/* ======================================================================== */
/* FP64 direct sum                                                           */
/* ======================================================================== */

void orc_direct_f64(const float *x, const float *y, const float *z, const float *m, int n,
                    const int32_t *targets, int n_targets, double softening, double *ax,
                    double *ay, double *az, int nthreads)
{
    const double e2 = softening * softening;
    int nt = targets ? n_targets : n;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
#endif
    for (int k = 0; k < nt; ++k) {
        int i = targets ? targets[k] : k;
        double xi = x[i], yi = y[i], zi = z[i];
        double sx = 0, sy = 0, sz = 0;
        for (int j = 0; j < n; ++j) {
            if (j == i) continue;
            double dx = (double)x[j] - xi, dy = (double)y[j] - yi, dz = (double)z[j] - zi;
            double r2 = dx * dx + dy * dy + dz * dz + e2;
            double inv = 1.0 / (r2 * sqrt(r2));
            double f = (double)m[j] * inv;
            sx += f * dx;
            sy += f * dy;
            sz += f * dz;
        }
        ax[k] = sx;
        ay[k] = sy;
        az[k] = sz;
    }
}

/* ======================================================================== */
/* MI355X design restatement                                                 */
/* ======================================================================== */

void orc_mt_root(const float *x, const float *y, const float *z, int n, float out[4])
{
    orc_tree tmp;
    memset(&tmp, 0, sizeof(tmp));
    orc_ref_bbox(&tmp, x, y, z, n);
    const float *bmin = tmp.bmin, *bmax = tmp.bmax;
    out[0] = (bmin[0] + bmax[0]) * 0.5f;
    out[1] = (bmin[1] + bmax[1]) * 0.5f;
    out[2] = (bmin[2] + bmax[2]) * 0.5f;
    out[3] = fmaxf(fmaxf(bmax[0] - bmin[0], bmax[1] - bmin[1]), bmax[2] - bmin[2]) * 0.5f *
             1.001f;
}

void orc_mt_keys(const float *x, const float *y, const float *z, int n, const float root[4],
                 uint64_t *keys)
{
    for (int i = 0; i < n; ++i) {
        float cx = root[0], cy = root[1], cz = root[2], h = root[3];
        const float px = x[i], py = y[i], pz = z[i];
        uint64_t key = 0;
        for (int l = 0; l < 21; ++l) {
            int ox = px > cx, oy = py > cy, oz = pz > cz;
            key = (key << 3) | (uint64_t)(ox | (oy << 1) | (oz << 2));
            h = h * 0.5f;
            cx = cx + (ox ? h : -h);
            cy = cy + (oy ? h : -h);
            cz = cz + (oz ? h : -h);
        }
        keys[i] = key;
    }
}

typedef struct { uint64_t key; int32_t idx; } orc_keyrec;

static int cmp_keyrec(const void *a, const void *b)
{
    const orc_keyrec *p = (const orc_keyrec *)a, *q = (const orc_keyrec *)b;
    if (p->key != q->key) return p->key < q->key ? -1 : 1;
    return (p->idx > q->idx) - (p->idx < q->idx);
}

void orc_mt_sort(const uint64_t *keys, int n, int32_t *perm)
{
    orc_keyrec *r = (orc_keyrec *)malloc((size_t)n * sizeof(orc_keyrec));
    for (int i = 0; i < n; ++i) { r[i].key = keys[i]; r[i].idx = i; }
    qsort(r, (size_t)n, sizeof(orc_keyrec), cmp_keyrec);
    for (int i = 0; i < n; ++i) perm[i] = r[i].idx;
    free(r);
}


/* Compressed octree over sorted positions: a node is a maximal run of >= 2
 * bodies that share `level` leading digits and differ at digit `level`
 * (level 21 = identical keys: every body is its own child). */
typedef struct {
    float cx, cy, cz, m, h2;
    int first, count; /* children are mt_child[first .. first+count) */
    int lo, hi;       /* sorted positions covered: [lo, hi] */
    double q[6];      /* traceless quadrupole about the COM: xx yy zz xy xz yz */
} mt_node;

/* Squared distance of the compressed-octree opening test:
 * ((dx dx) + dy dy) + dz dz with each addition fused into its product,
 * evaluated by fmaf (correctly rounded), so the CPU and GPU walks of the
 * same tree take the same decisions. */
static inline float mt_dist2(float dx, float dy, float dz)
{
    return fmaf(dz, dz, fmaf(dy, dy, dx * dx));
}

/* Opening test of the compressed walk: h^2 theta^-2 < r^2 with theta^-2 =
 * 1 / (theta theta) rounded once (the reference's h / r < theta,
 * BHTree.cu:576-577, without the square root and the division; for theta
 * a power of two, e.g. 0.5 or 0.25, both scalings are exact). */
static inline float mt_inv_theta2(float theta) { return 1.0f / (theta * theta); }
static inline int mt_accept(float h2, float ith2, float r2) { return h2 * ith2 < r2; }

typedef struct {
    const uint64_t *keys;
    const float *sx, *sy, *sz, *sm; /* bodies in sorted order */
    mt_node *nodes;
    int *child; /* >= 0: node id;  < 0: ~(sorted body position) */
    int n_nodes, n_child;
    float rhs;
} mt_tree;

static int lcp_digits(uint64_t a, uint64_t b)
{
    uint64_t d = a ^ b;
    if (!d) return 21;
    return (__builtin_clzll(d) - 1) / 3; /* bit 63 of a key is always zero */
}

static int digit_at(uint64_t key, int level) { return (int)((key >> (3 * (20 - level))) & 7); }

static int mt_build_rec(mt_tree *T, int a, int b)
{
    const int id = T->n_nodes++;
    const int level = lcp_digits(T->keys[a], T->keys[b - 1]);
    int nruns = 0;
    for (int s = a; s < b;) {
        int e = s + 1;
        if (level < 21)
            while (e < b && digit_at(T->keys[e], level) == digit_at(T->keys[s], level)) ++e;
        ++nruns;
        s = e;
    }
    const int first = T->n_child;
    T->n_child += nruns;
    double m = 0, cx = 0, cy = 0, cz = 0;
    int k = 0;
    for (int s = a; s < b; ++k) {
        int e = s + 1;
        if (level < 21)
            while (e < b && digit_at(T->keys[e], level) == digit_at(T->keys[s], level)) ++e;
        if (e - s == 1) {
            T->child[first + k] = ~s;
            m += T->sm[s];
            cx += (double)T->sm[s] * T->sx[s];
            cy += (double)T->sm[s] * T->sy[s];
            cz += (double)T->sm[s] * T->sz[s];
        } else {
            int c = mt_build_rec(T, s, e);
            T->child[first + k] = c;
            const mt_node *cn = &T->nodes[c];
            m += cn->m;
            cx += (double)cn->m * cn->cx;
            cy += (double)cn->m * cn->cy;
            cz += (double)cn->m * cn->cz;
        }
        s = e;
    }
    mt_node *nd = &T->nodes[id];
    nd->m = (float)m;
    nd->cx = m > 0 ? (float)(cx / m) : 0.f;
    nd->cy = m > 0 ? (float)(cy / m) : 0.f;
    nd->cz = m > 0 ? (float)(cz / m) : 0.f;
    /* parallel-axis sum of the children's quadrupoles about this COM */
    for (int q = 0; q < 6; ++q) nd->q[q] = 0;
    for (int c = 0; c < nruns; ++c) {
        int ch = T->child[first + c];
        double cm, d[3];
        const double *cq = NULL;
        if (ch < 0) {
            int b = ~ch;
            cm = T->sm[b];
            d[0] = (double)T->sx[b] - nd->cx; d[1] = (double)T->sy[b] - nd->cy;
            d[2] = (double)T->sz[b] - nd->cz;
        } else {
            const mt_node *cn = &T->nodes[ch];
            cm = cn->m;
            d[0] = (double)cn->cx - nd->cx; d[1] = (double)cn->cy - nd->cy;
            d[2] = (double)cn->cz - nd->cz;
            cq = cn->q;
        }
        double d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        nd->q[0] += cm * (3 * d[0] * d[0] - d2) + (cq ? cq[0] : 0);
        nd->q[1] += cm * (3 * d[1] * d[1] - d2) + (cq ? cq[1] : 0);
        nd->q[2] += cm * (3 * d[2] * d[2] - d2) + (cq ? cq[2] : 0);
        nd->q[3] += cm * 3 * d[0] * d[1] + (cq ? cq[3] : 0);
        nd->q[4] += cm * 3 * d[0] * d[2] + (cq ? cq[4] : 0);
        nd->q[5] += cm * 3 * d[1] * d[2] + (cq ? cq[5] : 0);
    }
    float h = ldexpf(T->rhs, -level);
    nd->h2 = h * h;
    nd->lo = a;
    nd->hi = b - 1;
    nd->first = first;
    nd->count = nruns;
    return id;
}

/* Quadrupole correction of an accepted node at R = COM - body (softened). */
static void mt_quad_term(const mt_node *nd, float dx, float dy, float dz, float rinv, float *fx,
                         float *fy, float *fz)
{
    float q0 = (float)nd->q[0], q1 = (float)nd->q[1], q2 = (float)nd->q[2];
    float q3 = (float)nd->q[3], q4 = (float)nd->q[4], q5 = (float)nd->q[5];
    float rinv2 = rinv * rinv;
    float rinv5 = rinv2 * rinv2 * rinv;
    float qx = q0 * dx + q3 * dy + q4 * dz;
    float qy = q3 * dx + q1 * dy + q5 * dz;
    float qz = q4 * dx + q5 * dy + q2 * dz;
    float rqr = dx * qx + dy * qy + dz * qz;
    float g = 2.5f * rqr * rinv5 * rinv2;
    *fx += g * dx - rinv5 * qx;
    *fy += g * dy - rinv5 * qy;
    *fz += g * dz - rinv5 * qz;
}

int orc_mt_force(const float *x, const float *y, const float *z, const float *m, int n,
                 float theta, float softening, int quad, float *self_ax, float *self_ay,
                 float *self_az, int32_t *n_acc, int32_t *n_open, int nthreads)
{
    return orc_mt_force_at(x, y, z, m, n, theta, softening, quad, NULL, n, self_ax, self_ay,
                           self_az, n_acc, n_open, nthreads);
}

int orc_mt_force_at(const float *x, const float *y, const float *z, const float *m, int n,
                    float theta, float softening, int quad, const int32_t *targets, int nt,
                    float *self_ax, float *self_ay, float *self_az, int32_t *n_acc,
                    int32_t *n_open, int nthreads)
{
    if (n <= 0) return 0;
    float root[4];
    orc_mt_root(x, y, z, n, root);
    uint64_t *keys = (uint64_t *)malloc((size_t)n * sizeof(uint64_t));
    uint64_t *skeys = (uint64_t *)malloc((size_t)n * sizeof(uint64_t));
    int32_t *perm = (int32_t *)malloc((size_t)n * sizeof(int32_t));
    float *sb = (float *)malloc((size_t)n * 4 * sizeof(float));
    orc_mt_keys(x, y, z, n, root, keys);
    orc_mt_sort(keys, n, perm);
    int32_t *inv = (int32_t *)malloc((size_t)n * sizeof(int32_t));
    for (int q = 0; q < n; ++q) inv[perm[q]] = q;
    for (int s = 0; s < n; ++s) {
        skeys[s] = keys[perm[s]];
        sb[s] = x[perm[s]];
        sb[n + s] = y[perm[s]];
        sb[2 * n + s] = z[perm[s]];
        sb[3 * n + s] = m[perm[s]];
    }
    mt_tree T;
    memset(&T, 0, sizeof(T));
    T.keys = skeys;
    T.sx = sb; T.sy = sb + n; T.sz = sb + 2 * n; T.sm = sb + 3 * n;
    T.nodes = (mt_node *)malloc((size_t)(n > 1 ? n : 1) * sizeof(mt_node));
    T.child = (int *)malloc((size_t)2 * (n > 1 ? n : 1) * sizeof(int));
    T.rhs = root[3];
    if (n > 1) mt_build_rec(&T, 0, n);
    const float ith2 = mt_inv_theta2(theta), eps2 = softening * softening;
    int overflow = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : overflow) num_threads(nthreads)
#endif
    for (int t = 0; t < nt; ++t) {
        /* target t: body targets[t] (all bodies in original order without a list) */
        const int s = inv[targets ? targets[t] : t];
        const float bx = T.sx[s], by = T.sy[s], bz = T.sz[s];
        float fx = 0.f, fy = 0.f, fz = 0.f;
        int acc = 0, opn = 0;
        if (n > 1) {
            int stack[512];
            int top = 0;
            /* the root is tested like any other node */
            {
                const mt_node *r = &T.nodes[0];
                float dx = r->cx - bx, dy = r->cy - by, dz = r->cz - bz;
                float r2 = mt_dist2(dx, dy, dz);
                if (mt_accept(r->h2, ith2, r2)) {
                    float rinv = 1.0f / sqrtf(r2 + eps2);
                    float f = r->m * rinv * rinv * rinv;
                    fx += f * dx; fy += f * dy; fz += f * dz;
                    if (quad) mt_quad_term(r, dx, dy, dz, rinv, &fx, &fy, &fz);
                    ++acc;
                } else {
                    ++opn;
                    stack[top++] = 0;
                }
            }
            while (top > 0) {
                const mt_node *p = &T.nodes[stack[--top]];
                for (int k = 0; k < p->count; ++k) {
                    int c = T.child[p->first + k];
                    float cxn, cyn, czn, cm;
                    int leaf = c < 0;
                    if (leaf) {
                        int b = ~c;
                        if (b == s) continue;
                        cxn = T.sx[b]; cyn = T.sy[b]; czn = T.sz[b]; cm = T.sm[b];
                    } else {
                        cxn = T.nodes[c].cx; cyn = T.nodes[c].cy; czn = T.nodes[c].cz;
                        cm = T.nodes[c].m;
                    }
                    float dx = cxn - bx, dy = cyn - by, dz = czn - bz;
                    float r2 = mt_dist2(dx, dy, dz);
                    if (leaf || mt_accept(T.nodes[c].h2, ith2, r2)) {
                        float rinv = 1.0f / sqrtf(r2 + eps2);
                        float f = cm * rinv * rinv * rinv;
                        fx += f * dx; fy += f * dy; fz += f * dz;
                        if (quad && !leaf) mt_quad_term(&T.nodes[c], dx, dy, dz, rinv, &fx, &fy, &fz);
                        ++acc;
                    } else {
                        ++opn;
                        if (top < 512) stack[top++] = c; else ++overflow;
                    }
                }
            }
        }
        if (self_ax) { self_ax[t] = fx; self_ay[t] = fy; self_az[t] = fz; }
        if (n_acc) n_acc[t] = acc;
        if (n_open) n_open[t] = opn;
    }
    free(keys); free(skeys); free(perm); free(inv); free(sb); free(T.nodes); free(T.child);
    return overflow;
}

/* Interaction lists of the monopole walk of orc_mt_force: for every target
 * (original order) the accepted nodes as sorted-position ranges [lo, hi]
 * (a leaf has lo == hi; the target's own leaf is skipped).  offsets[n + 1]
 * is filled always; lo / hi (capacity `cap`) only when both are non-NULL.
 * perm[s] = original index of sorted position s.  Returns the total number
 * of accepted nodes, or -1 when cap is too small. */
long orc_mt_lists(const float *x, const float *y, const float *z, const float *m, int n,
                  float theta, int32_t *perm, int64_t *offsets, int32_t *lo, int32_t *hi,
                  long cap)
{
    if (n <= 0) return 0;
    float root[4];
    orc_mt_root(x, y, z, n, root);
    uint64_t *keys = (uint64_t *)malloc((size_t)n * sizeof(uint64_t));
    uint64_t *skeys = (uint64_t *)malloc((size_t)n * sizeof(uint64_t));
    float *sb = (float *)malloc((size_t)n * 4 * sizeof(float));
    orc_mt_keys(x, y, z, n, root, keys);
    orc_mt_sort(keys, n, perm);
    for (int s = 0; s < n; ++s) {
        skeys[s] = keys[perm[s]];
        sb[s] = x[perm[s]];
        sb[n + s] = y[perm[s]];
        sb[2 * n + s] = z[perm[s]];
        sb[3 * n + s] = m[perm[s]];
    }
    mt_tree T;
    memset(&T, 0, sizeof(T));
    T.keys = skeys;
    T.sx = sb; T.sy = sb + n; T.sz = sb + 2 * n; T.sm = sb + 3 * n;
    T.nodes = (mt_node *)malloc((size_t)(n > 1 ? n : 1) * sizeof(mt_node));
    T.child = (int *)malloc((size_t)2 * (n > 1 ? n : 1) * sizeof(int));
    T.rhs = root[3];
    if (n > 1) mt_build_rec(&T, 0, n);
    const float ith2 = mt_inv_theta2(theta);
    int32_t *cnt = (int32_t *)calloc((size_t)n, sizeof(int32_t));
    const int fill = lo && hi;
    long total = 0;
    for (int pass = 0; pass < (fill ? 2 : 1); ++pass) {
        if (pass == 1) {
            offsets[0] = 0;
            for (int i = 0; i < n; ++i) offsets[i + 1] = offsets[i] + cnt[i];
            total = offsets[n];
            if (total > cap) { total = -1; break; }
        }
        for (int s = 0; s < n && n > 1; ++s) {
            const int i = perm[s];
            long w = pass == 1 ? offsets[i] : 0;
            int c_i = 0;
            const float bx = T.sx[s], by = T.sy[s], bz = T.sz[s];
            int stack[512];
            int top = 0;
            const mt_node *r = &T.nodes[0];
            float dx = r->cx - bx, dy = r->cy - by, dz = r->cz - bz;
            if (mt_accept(r->h2, ith2, mt_dist2(dx, dy, dz))) {
                if (pass == 1) { lo[w] = r->lo; hi[w] = r->hi; ++w; }
                ++c_i;
            } else {
                stack[top++] = 0;
            }
            while (top > 0) {
                const mt_node *p = &T.nodes[stack[--top]];
                for (int k = 0; k < p->count; ++k) {
                    const int c = T.child[p->first + k];
                    if (c < 0) {
                        const int b = ~c;
                        if (b == s) continue;
                        if (pass == 1) { lo[w] = b; hi[w] = b; ++w; }
                        ++c_i;
                        continue;
                    }
                    const mt_node *nd = &T.nodes[c];
                    dx = nd->cx - bx; dy = nd->cy - by; dz = nd->cz - bz;
                    if (mt_accept(nd->h2, ith2, mt_dist2(dx, dy, dz))) {
                        if (pass == 1) { lo[w] = nd->lo; hi[w] = nd->hi; ++w; }
                        ++c_i;
                    } else if (top < 512) {
                        stack[top++] = c;
                    }
                }
            }
            cnt[i] = c_i;
        }
        if (pass == 0 && !fill) {
            offsets[0] = 0;
            for (int i = 0; i < n; ++i) offsets[i + 1] = offsets[i] + cnt[i];
            total = offsets[n];
        }
    }
    free(cnt); free(keys); free(skeys); free(sb); free(T.nodes); free(T.child);
    return total;
}

/* ======================================================================== */
/* Whole reference step                                                      */
/* ======================================================================== */

/* main.cu:113-119 + :123-127: rebuild the tree and overwrite a, then add the
 * perturber pull at sim_time. */
static int orc_forces(orc_tree *t, const float *x, const float *y, const float *z,
                      float *ax, float *ay, float *az, const float *m, int n,
                      const orc_step_params *p, const orc_perturber *pert, int nthreads)
{
    orc_ref_bbox(t, x, y, z, n);
    int lost = orc_ref_build(t, x, y, z, m, n);
    orc_ref_summarize(t);
    orc_ref_force(t, x, y, z, n, p->theta, p->softening, p->star_mass, ax, ay, az, NULL,
                  NULL, NULL, NULL, NULL, nthreads);
    if (pert->enabled)
        orc_perturber_force(ax, ay, az, x, y, z, pert->x, pert->y, pert->z, pert->mass_msun,
                            p->softening, n);
    return lost;
}

int orc_ref_init_forces(float *x, float *y, float *z, float *vx, float *vy, float *vz,
                        float *ax, float *ay, float *az, const float *m, int n,
                        const orc_step_params *p, int nthreads)
{
    (void)vx; (void)vy; (void)vz;
    orc_tree *t = orc_tree_new(n);
    orc_perturber pert;
    orc_make_perturber(&pert, p->perturber_enabled, p->pert_mass_mjup, p->pert_radius_au,
                       p->pert_phase0, p->star_mass);
    orc_update_perturber(&pert, 0.0f);
    int lost = orc_forces(t, x, y, z, ax, ay, az, m, n, p, &pert, nthreads);
    orc_tree_free(t);
    return lost;
}

/* main.cu:101-140 */
int orc_ref_steps(float *x, float *y, float *z, float *vx, float *vy, float *vz, float *ax,
                  float *ay, float *az, const float *m, int n, const orc_step_params *p,
                  int step0, int steps, int nthreads)
{
    orc_tree *t = orc_tree_new(n);
    orc_perturber pert;
    orc_make_perturber(&pert, p->perturber_enabled, p->pert_mass_mjup, p->pert_radius_au,
                       p->pert_phase0, p->star_mass);
    int lost = 0;
    for (int step = step0; step < step0 + steps; ++step) {
        orc_half_kick(vx, vy, vz, ax, ay, az, p->dt * 0.5f, n);
        orc_drift(x, y, z, vx, vy, vz, p->dt, n);
        float sim_time_now = (float)(step + 1) * p->dt;
        orc_update_perturber(&pert, sim_time_now);
        lost += orc_forces(t, x, y, z, ax, ay, az, m, n, p, &pert, nthreads);
        if (p->collisions)
            orc_ref_collide(t, x, y, z, vx, vy, vz, n, p->r_coll, p->restitution,
                            p->exact_collide, nthreads);
        orc_half_kick(vx, vy, vz, ax, ay, az, p->dt * 0.5f, n);
        orc_gas_drag(x, y, z, vx, vy, vz, p->gas_drag_rate, p->dt, p->star_mass, n);
    }
    orc_tree_free(t);
    return lost;
}
