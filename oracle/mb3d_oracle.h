/* mb3d_oracle.h -- TEST INFRASTRUCTURE, not product code.
 *
 * The f64 CPU oracle: a restatement of mb3d-rs's CPU ray marcher
 * (src/calc.rs, src/iteration.rs, src/formulas.rs, src/math.rs,
 * src/deao.rs) for the built-in formulas and their alternating 3D and 4D
 * hybrids, with Julia mode, and of its painter (src/lighting.rs).  Only the tests, the smoke check and the CPU
 * baseline leg of the benchmark load it, always as the checker.
 *
 * It shares the plain data layouts of include/mb3d_hip.h (scene,
 * parameters, post job), and keeps its own restatement of the parameter
 * set-up (CalcParams::new), so that the product's set-up is checked
 * against it field by field.
 */
#ifndef MB3D_ORACLE_H
#define MB3D_ORACLE_H

#include <stdint.h>
#include "../include/mb3d_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Work counters of an oracle run (same meaning as mb3d_hip_counters). */
typedef struct orc_counters {
    uint64_t pixels, hits, steps, de_evals;
    uint64_t its_p8, its_box, its_other;  /* Integer Power, Amazing Box, the other formulas */
} orc_counters;

/* Oracle output per pixel of a rectangle; any pointer may be NULL. */
typedef struct orc_out {
    double *depth;      /* mzz (step widths); +inf background; 0 inside at the start */
    uint16_t *iters;    /* it_result after the bisection (0 background) */
    double *normals;    /* 3 per pixel, unit, view (vgrads) frame */
    uint32_t *gbuf;     /* 5 words per pixel (the card record) */
    int32_t *steps;     /* march steps per pixel */
} orc_out;

const char *orc_last_error(void);

/* CalcParams::new restated (src/calc.rs:375-898), built-in formulas. */
int32_t orc_params_from_scene(const mb3d_hip_scene *scene, mb3d_hip_params *params);

/* pixel_seed (src/calc.rs:1125-1140). */
int32_t orc_pixel_seed(int64_t y, int64_t x);

/* One formula iteration (Formula::iterate, src/formulas.rs:576-641);
 * `rout` is the stored squared radius (Real Power, Bulbox), `j` the
 * constant (Aexion C turns it). */
void orc_iterate(const mb3d_hip_formula *f, double v[4], double j[4], double rout, int32_t ade);

/* de_at (src/calc.rs:2695-2703): the DE at a world point in world units,
 * and the iteration count there. */
double orc_de_at(const mb3d_hip_params *p, const double pos[3], int32_t *its);

/* pixel_ray (src/calc.rs:2324-2355 via Marcher::pixel_ray): start and
 * direction (one step long) of pixel (x, y). */
void orc_pixel_ray(const mb3d_hip_params *p, int32_t x, int32_t y, double c[3], double v[3]);

/* The G-buffer of rows [row0, row0+rows) by the CPU marcher with row
 * interleaved threads (src/render.rs:757-803), then the post job on it
 * (per-pixel hard shadows, DE AO).  threads <= 0: all cores. */
int32_t orc_render(const mb3d_hip_params *p, const mb3d_hip_post_job *post, int32_t row0,
                   int32_t rows, const orc_out *out, int32_t threads, orc_counters *cnt);

/* The cone march (src/cone.rs, src/calc.rs:2357-2435) with one level of
 * blocks of bw x bh pixels: each block's safe depth (+inf: background) and
 * the dynamic fog's count up to it, block rows of (width + bw - 1) / bw. */
int32_t orc_cone_grid(const mb3d_hip_params *p, int32_t bw, int32_t bh, double *t, float *count,
                      int32_t threads);

/* orc_render with the cone grid: mode 1, a block whose cone passed the end
 * is background; mode 2, also every pixel's march from its block's depth. */
int32_t orc_render_cone(const mb3d_hip_params *p, int32_t bw, int32_t bh, int32_t mode, int32_t row0,
                        int32_t rows, const orc_out *out, int32_t threads, orc_counters *cnt);

/* The post job alone on a G-buffer (5 words per pixel, in and out). */
int32_t orc_post(const mb3d_hip_params *p, const mb3d_hip_post_job *post, int32_t row0,
                 int32_t rows, uint32_t *gbuf, int32_t threads);

/* One pixel's DE AO (deao_row, src/deao.rs:268-418) on its card record,
 * with what it computed, for the tests: out (174 doubles) = the rebuilt
 * surface point, each ray's world step (33 x 3), each ray's openness times
 * de_mul (33), then step_ao, max_dist, the rays' DE stop, md_d10, de_mul,
 * d_step_mul, then each ray's first step factor (33).  Returns the ray
 * count (0: not a surface pixel); *amb the amb_shadow word.  `rnd`: the
 * row generator's state before this pixel, advanced by its draws (needed
 * with the random first step, else may be NULL). */
int32_t orc_deao_probe(const mb3d_hip_params *p, const mb3d_hip_post_job *job, int32_t x, int32_t y,
                       const uint32_t rec[5], double *out, int32_t *amb, int32_t *rnd);

/* The DE AO's row generator (DeaoRows::row, get_rand, src/deao.rs:262-270):
 * the seed of picture row y; n draws from `seed`: the states and numbers. */
int32_t orc_deao_row_seed(int32_t y);
void orc_deao_rand(int32_t seed, int32_t n, int32_t *states, double *vals);

/* create_xy_vecs_from_normals (src/math.rs:218-232): two vectors across
 * the normal n, for the smoothing sweeps of calculate_normals. */
void orc_xy_vecs(const double n[3], double vx[3], double vy[3]);

/* The march of one ray from `origin` with direction `dir` (one step long)
 * at depth `zz`, as Marcher::march_ray_from (src/calc.rs:2519-2531): the
 * depth reached (+inf: background), its iteration count and steps. */
double orc_march_ray(const mb3d_hip_params *p, const double origin[3], const double dir[3],
                     double zz, int32_t *its, int32_t *steps);

/* SrTracer::march_ray (src/reflect.rs:1464-1674) for each task: the card's
 * reflection batch (no volumetric light or ground), with calc_hs and
 * deao_1pos at the hits when `opts` asks (src/reflect.rs:805-1090).  Tasks
 * in chunks of 64 over threads (<= 0: all cores). */
int32_t orc_march_rays(const mb3d_hip_params *p, const mb3d_hip_ray_task *tasks, int64_t n, int32_t depth_sqrt,
                       const mb3d_hip_ray_opts *opts, mb3d_hip_ray_hit *hits, int32_t threads);

/* the DE (in steps) and iteration count at n points (3 f64 each), as
 * orc_de_at without its step-width scale */
int32_t orc_de_many(const mb3d_hip_params *p, const double *pos, int64_t n, double *de, int32_t *its, int32_t threads);

/* the Monte Carlo renderer's marches (McMarch::run, src/mc.rs:488-977), as
 * mb3d_hip_mc_batch */
int32_t orc_mc_batch(const mb3d_hip_params *p, const mb3d_hip_mc_task *tasks, int64_t n, int32_t secant,
                     mb3d_hip_mc_res *res, int32_t threads);

/* One pixel's march with what do_color saw at its surface (tests): out =
 * the orbit's end v (3), the surface point c (3), rout, rold; words = the
 * otrap and si_gradient words.  1 at a surface, 0 otherwise. */
int32_t orc_color_probe(const mb3d_hip_params *params, int32_t x, int32_t y, double *out, uint16_t *words);

/* ---- the painter (paint_oracle.cpp) ---- */

/* Lighting::default (src/lighting.rs:290-392). */
void orc_lighting_default(mb3d_hip_lighting *lighting);

/* paint_camera and light_vals (src/render.rs:2340-2380, 2285-2319) for the
 * common optic without stereo. */
int32_t orc_paint_setup(const mb3d_hip_scene *scene, const mb3d_hip_params *params,
                        const mb3d_hip_lighting *lighting, mb3d_hip_paint_camera *camera,
                        mb3d_hip_light_vals *light_vals);

/* gbuffer::pack / unpack (src/gbuffer.rs:60-98) between card records (5
 * words) and 18-byte file records. */
void orc_gbuf_pack(const uint32_t *gbuf, int64_t n, uint8_t *out);
void orc_gbuf_unpack(const uint8_t *in, int64_t n, uint32_t *gbuf);

/* GetCosTabVal (src/lighting.rs:501-541), and the 8 x 128 tables of cos_tabs. */
float orc_cos_tab_val(int32_t tnr, float dotp, float rough, int32_t sqr);
void orc_cos_tables(float *out);

/* pixel_color_s (src/lighting.rs:1739-1760) for every record of the
 * rectangle (x0, y0, w, h): 3 channels a pixel in `format`
 * (MB3D_HIP_PAINT_*), rows interleaved over threads (<= 0: all cores). */
int32_t orc_paint(const mb3d_hip_paint_camera *camera, const mb3d_hip_light_vals *light_vals,
                  const uint32_t *gbuf, int32_t x0, int32_t y0, int32_t w, int32_t h, int32_t format,
                  void *rgb, int32_t threads);

#ifdef __cplusplus
}
#endif

#endif
