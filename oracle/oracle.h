/*
 * oracle.h -- CPU restatement of the reference Barnes-Hut disk step.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product path (the HIP library,
 * its C-ABI or the C++/Python host mirrors) may link, load or call this code.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it,
 * and only as the checker or as the timed CPU baseline.
 *
 * Two families of functions live here:
 *
 *  orc_ref_*   restate the reference algorithm (ualiberik/nbody-sim) line by
 *              line in plain C with float arithmetic and no FMA contraction.
 *              Each function cites the reference file:line it follows.
 *
 *  orc_mt_*    restate the MI355X design (Morton keys by float descent,
 *              stable key sort, compressed octree from sorted keys, the same
 *              per-body opening rule).  The GPU kernels must reproduce the
 *              integer outputs (keys, order, FoF labels) bit for bit.
 *
 *  orc_direct_f64  O(N^2) direct sum in double: the accuracy anchor.
 *
 * Pinning (which check fixes which function) is listed per function in
 * tests/test_oracle_*.py; functions without an independent check say
 * "parity unpinned" in their comment.
 */
#ifndef NBX_ORACLE_H
#define NBX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Reference canonical octree (BHTree.cu)                                    */
/* ------------------------------------------------------------------------ */

typedef struct orc_tree {
    int    n_bodies;     /* leaves [0,n) are bodies                         */
    int    n_total;      /* n + 8n node pool, root = n_total-1 (BHTree.cuh:4) */
    int    next_node;    /* allocator (BHTree.cu:245)                        */
    float *pos_x, *pos_y, *pos_z, *mass, *cell_size;
    int   *child;        /* [n_total*8], -1 empty                            */
    int   *count;
    float  bmin[3], bmax[3];
    int    lost;         /* bodies dropped by max_iter / pool overflow (F6b) */
    int    max_depth;    /* deepest level reached                            */
} orc_tree;

orc_tree *orc_tree_new(int n_bodies);
void      orc_tree_free(orc_tree *t);

/* BHTree.cu:76-134 + 422-452: AABB then 1.001 padding about the midpoint. */
void orc_ref_bbox(orc_tree *t, const float *x, const float *y, const float *z, int n);

/* BHTree.cu:147-159 + 175-277 + 454-499: canonical single-body-leaf octree
 * built by sequential insertion (same structure as the reference's race-free
 * result, node indices differ).  Returns the number of bodies lost. */
int  orc_ref_build(orc_tree *t, const float *x, const float *y, const float *z,
                   const float *m, int n);

/* BHTree.cu:282-317: COM = sum(m*x)/sum(m) over children k=0..7, bottom-up.
 * Full reduction (no 20-pass cap, see finding F5). */
void orc_ref_summarize(orc_tree *t);

/* BHTree.cu:322-352: DFS leaf order (children pushed 7..0, popped 0..7). */
int  orc_ref_sort(const orc_tree *t, int32_t *sort_idx);

/* BHTree.cu:530-598: per-body walk.  Star term first, then accepted nodes in
 * stack order, accumulated into one float accumulator (reference order).
 * Optional outputs (may be NULL):
 *   self_a*  : self-gravity only, accumulated from 0 in the same order
 *   n_acc    : accepted interactions per body (leaves + nodes, self excluded)
 *   n_open   : opened internal nodes per body
 * Returns the number of stack overflows (F6a). */
int  orc_ref_force(const orc_tree *t, const float *x, const float *y, const float *z,
                   int n, float theta, float softening, float star_mass,
                   float *ax, float *ay, float *az,
                   float *self_ax, float *self_ay, float *self_az,
                   int32_t *n_acc, int32_t *n_open, int nthreads);

/* BHTree.cu:620-708: two-pass collision impulse.
 * exact=0 : the reference pruning |COM-x| - 1.732 h < r_coll (finding F7);
 * exact=1 : exact pair set dist^2 < r_coll^2 (prune with 2*sqrt(3)*h).
 * dv accumulates in traversal order; v += dv applied at the end (:676-686).
 * Returns the number of colliding (i,j) ordered pairs found. */
long orc_ref_collide(const orc_tree *t, const float *x, const float *y, const float *z,
                     float *vx, float *vy, float *vz, int n,
                     float r_coll, float restitution, int exact, int nthreads);

/* ------------------------------------------------------------------------ */
/* Integrator / perturber (Integrator.cu, Perturber.cu, Perturber.h)          */
/* ------------------------------------------------------------------------ */
void orc_half_kick(float *vx, float *vy, float *vz, const float *ax, const float *ay,
                   const float *az, float dt_half, int n);              /* :7-15  */
void orc_drift(float *x, float *y, float *z, const float *vx, const float *vy,
               const float *vz, float dt, int n);                       /* :17-25 */
void orc_gas_drag(const float *x, const float *y, const float *z, float *vx, float *vy,
                  float *vz, float drag_rate, float dt, float star_mass, int n); /* :40-76,100-109 */
void orc_perturber_force(float *ax, float *ay, float *az, const float *x, const float *y,
                         const float *z, float px, float py, float pz, float pmass,
                         float softening, int n);                       /* Perturber.cu:13-35 */

typedef struct orc_perturber {
    int   enabled;
    float r, mass_msun, phase0, omega, x, y, z;
} orc_perturber;
/* Perturber.h:39-57 and :60-66 */
void orc_make_perturber(orc_perturber *p, int enabled, float mass_mjup, float radius_au,
                        float phase0, float star_mass);
void orc_update_perturber(orc_perturber *p, float sim_time);

/* DiskInit.cpp:10-16 */
float orc_particle_physical_radius(float particle_mass_msun);

/* ------------------------------------------------------------------------ */
/* FoF (Aggregate.cpp:26-76)                                                 */
/* ------------------------------------------------------------------------ */
/* O(N^2) union-find on the float predicate (dx^2+dy^2)+dz^2 < ll^2.
 * labels[i] = smallest particle index of i's component (canonical form of
 * the reference's partition, finding F8).  Returns the number of groups of
 * size >= 2. */
int orc_fof_n2(const float *x, const float *y, const float *z, int n, float ll,
               int32_t *labels);
/* Same partition through a uniform cell grid (for large n); identical result. */
int orc_fof_grid(const float *x, const float *y, const float *z, int n, float ll,
                 int32_t *labels);

/* ------------------------------------------------------------------------ */
/* FP64 direct sum                                                            */
/* ------------------------------------------------------------------------ */
/* Self-gravity on the listed targets (all if targets==NULL) in double with
 * Plummer softening, excluding i==j. */
void orc_direct_f64(const float *x, const float *y, const float *z, const float *m, int n,
                    const int32_t *targets, int n_targets, double softening,
                    double *ax, double *ay, double *az, int nthreads);

/* ------------------------------------------------------------------------ */
/* MI355X design restatement                                                  */
/* ------------------------------------------------------------------------ */
/* Root cube exactly as BHTree.cu:441-446 then :187-191 (float ops, in order):
 * out = {cx, cy, cz, half}. */
void orc_mt_root(const float *x, const float *y, const float *z, int n, float out[4]);

/* 63-bit key: 21 octant digits from the float descent of BHTree.cu:253-256,
 * digit = (x>cx) | (y>cy)<<1 | (z>cz)<<2, first level most significant. */
void orc_mt_keys(const float *x, const float *y, const float *z, int n, const float root[4],
                 uint64_t *keys);

/* Stable sort by key (ties by index): perm[s] = body at sorted position s. */
void orc_mt_sort(const uint64_t *keys, int n, int32_t *perm);

/* Compressed octree of the sorted keys and the per-body walk with the
 * reference opening rule written as h^2 theta^-2 < r^2.  Tree terms are
 * accumulated from zero (self-gravity), the star term is returned apart.
 * n_acc / n_open as in orc_ref_force (n_open counts compressed nodes).
 * quad=1 adds the traceless quadrupole of each accepted internal node. */
int orc_mt_force(const float *x, const float *y, const float *z, const float *m, int n,
                 float theta, float softening, int quad,
                 float *self_ax, float *self_ay, float *self_az,
                 int32_t *n_acc, int32_t *n_open, int nthreads);
/* The same walk for the bodies targets[0..nt) only (original indices; the
 * tree holds all n bodies); outputs are indexed by t < nt.  targets = NULL
 * with nt = n is orc_mt_force. */
int orc_mt_force_at(const float *x, const float *y, const float *z, const float *m, int n,
                    float theta, float softening, int quad, const int32_t *targets, int nt,
                    float *self_ax, float *self_ay, float *self_az,
                    int32_t *n_acc, int32_t *n_open, int nthreads);

/* Interaction lists of the same monopole walk: per target (original order)
 * the accepted nodes as sorted-position ranges [lo, hi] (leaf: lo == hi).
 * offsets[n+1] always; lo/hi (capacity cap) when both non-NULL; perm[n] =
 * sorted -> original.  Returns the total, or -1 if cap is too small. */
long orc_mt_lists(const float *x, const float *y, const float *z, const float *m, int n,
                  float theta, int32_t *perm, int64_t *offsets, int32_t *lo, int32_t *hi,
                  long cap);

/* ------------------------------------------------------------------------ */
/* Whole reference step (main.cu:101-140), for the CPU baseline               */
/* ------------------------------------------------------------------------ */
typedef struct orc_step_params {
    float dt, theta, softening, star_mass, r_coll, restitution, gas_drag_rate;
    int   perturber_enabled;
    float pert_mass_mjup, pert_radius_au, pert_phase0;
    int   collisions;     /* 1 = run the collision pass                   */
    int   exact_collide;  /* 1 = exact pair set, 0 = reference pruning     */
} orc_step_params;

/* Initial forces (main.cu:85-96) when step0==0 and ax is to be filled,
 * then `steps` KDK steps starting at step index step0.  Returns lost bodies. */
int orc_ref_init_forces(float *x, float *y, float *z, float *vx, float *vy, float *vz,
                        float *ax, float *ay, float *az, const float *m, int n,
                        const orc_step_params *p, int nthreads);
int orc_ref_steps(float *x, float *y, float *z, float *vx, float *vy, float *vz,
                  float *ax, float *ay, float *az, const float *m, int n,
                  const orc_step_params *p, int step0, int steps, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
