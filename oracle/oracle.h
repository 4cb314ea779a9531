/*
 * CPU oracle for the vexel-hip hot path: a plain C/C++ restatement of the reference's
 * kernels, used ONLY as the checker by tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg. Nothing in the product library links or calls it.
 *
 * Every function follows the reference op for op in IEEE binary32, built with
 * -ffp-contract=off (the reference's IR never fuses a multiply-add), with the
 * reference file:line it restates next to it.  Reference paths are shortened:
 *   core/  = vexelray-sim-fluid-core/src/main/java/dev/vexelray/sim/fluid/
 *   test/  = vexelray-sim-fluid-core/src/test/java/dev/vexelray/sim/fluid/
 *
 * What pins each function (none of it is pinned by retyping its formula):
 *   - orc_swe_step: Ritter convergence, first-step size, clock end, closed-box mass,
 *     lake at rest (test/stencil/ShallowWaterTest.java), textbook HLL special cases.
 *   - orc_swe_diagnostics: closed forms of test/stencil/DiagnosticsTest.java.
 *   - orc_mpm3_* / orc_mpm2_*: lone-particle free fall (closed form), grid mass and
 *     momentum = particle sums, an independent FP64 splat (core/particle/ParticleSplat.java),
 *     dam-break bounds (test/particle/FlipTest.java, Flip3Test.java).
 *   - orc_sort_*: numpy's stable argsort and the reference's cellStarts contract.
 *   - orc_rng_*: published java.util.Random sequences (fixed known values).
 */
#ifndef VEXEL_ORACLE_H
#define VEXEL_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- java.util.Random (48-bit LCG), for the reference tests' seeded inputs ---- */
typedef struct orc_rng orc_rng;
orc_rng* orc_rng_new(int64_t seed);
void orc_rng_free(orc_rng* r);
int32_t orc_rng_next_int(orc_rng* r);
int32_t orc_rng_next_int_bound(orc_rng* r, int32_t bound);
float orc_rng_next_float(orc_rng* r);
double orc_rng_next_double(orc_rng* r);
double orc_rng_next_gaussian(orc_rng* r);
/* Draws `count` repetitions of `pattern` ('f' nextFloat, 'g' (float)nextGaussian,
 * 'd' (float)nextDouble) into out[count * strlen(pattern)], in call order. */
void orc_rng_pattern(orc_rng* r, const char* pattern, int64_t count, float* out);
/* The reference's Fisher-Yates shuffle order (test/particle/ScatterTest.java:481-493). */
void orc_rng_shuffle_order(orc_rng* r, int32_t n, int32_t* order);

/* ---- HLL shallow water, core/stencil/ShallowWater.java ---- */
/* edges: west, east, south, north; 0 = WALL, 1 = OPEN (core/stencil/Edge.java). */
void orc_swe_step(int nx, int ny, const int edges[4],
                  float* out_h, float* out_hu, float* out_hv,
                  const float* in_h, const float* in_hu, const float* in_hv,
                  const int32_t params[5],
                  const int32_t* speed_in, int32_t* speed_out, int32_t* speed_clear,
                  const int32_t* budget_in, int32_t* budget_out, int32_t* clamped);
/* The same step over `threads` host threads (OpenMP): bit-identical to orc_swe_step, as each
 * cell is computed alone and the speed maximum and clamp count do not depend on the order. */
void orc_swe_step_omp(int nx, int ny, const int edges[4],
                      float* out_h, float* out_hu, float* out_hv,
                      const float* in_h, const float* in_hu, const float* in_hv,
                      const int32_t params[5],
                      const int32_t* speed_in, int32_t* speed_out, int32_t* speed_clear,
                      const int32_t* budget_in, int32_t* budget_out, int32_t* clamped, int threads);
int32_t orc_swe_fastest_wave(int64_t cells, const float* h, const float* hu, const float* hv,
                             double g, double dry);
/* core/stencil/Diagnostics.java:84-122; out = volume, minDepth, maxDepth, maxSpeed,
 * maxFroude, maxCourant, nonFinite, negative, wet. */
void orc_swe_diagnostics(int64_t cells, const float* h, const float* hu, const float* hv,
                         double g, double dry, double dt, double dx, double dy, double out[9]);
/* FP64 twin of one step (same rule, double arithmetic, dt given in seconds). */
void orc_swe_step_f64(int nx, int ny, const int edges[4],
                      double* out_h, double* out_hu, double* out_hv,
                      const double* in_h, const double* in_hu, const double* in_hv,
                      double g, double dx, double dy, double dry, double dt);

/* ---- 3D MLS-MPM / APIC, core/particle/Flip3.java ----
 * Particle fields in the Flip3.PARTICLE order: x y z u v w m j c00..c22 (17).
 * p[f] points at field f's array of `count` floats. Grid: gm gmu gmv gmw, gu gv gw. */
void orc_mpm3_clear(int nx, int ny, int nz, float* gm, float* gmu, float* gmv, float* gmw);
void orc_mpm3_scatter(int nx, int ny, int nz, int64_t count, float* const p[17],
                      const int32_t params[7], float* gm, float* gmu, float* gmv, float* gmw);
/* The 3D fixed-point scatter: every amount orc_mpm3_scatter adds, rounded to a 64-bit integer
 * (mass in 2^-40, momentum in 2^-32, ties to even) and summed with wrap-around into q (zeroed
 * first); orc_mpm3_decode_fixed turns the sums into gm gmu gmv gmw. */
void orc_mpm3_scatter_fixed(int nx, int ny, int nz, int64_t count, float* const p[17], const int32_t params[7],
                            int64_t* qm, int64_t* qmu, int64_t* qmv, int64_t* qmw);
void orc_mpm3_decode_fixed(int64_t nodes, const int64_t* qm, const int64_t* qmu, const int64_t* qmv,
                           const int64_t* qmw, float* gm, float* gmu, float* gmv, float* gmw);
void orc_mpm3_grid(int nx, int ny, int nz, const int32_t params[7],
                   const float* gm, const float* gmu, const float* gmv, const float* gmw,
                   float* gu, float* gv, float* gw);
void orc_mpm3_advect(int nx, int ny, int nz, int64_t count, float* const p[17],
                     const int32_t params[7], const float* gu, const float* gv, const float* gw);
/* One whole step, as core/particle/Flip3Step.java:81-118 lists it. */
void orc_mpm3_step(int nx, int ny, int nz, int64_t count, float* const p[17],
                   const int32_t params[7], float* const grid[7]);
/* The same substep on `threads` host threads (OpenMP), for the CPU baseline only: its P2G adds
 * per-thread boxes of sums into the grid, so its sums are reassociated. */
void orc_mpm3_step_omp(int nx, int ny, int nz, int64_t count, float* const p[17],
                       const int32_t params[7], float* const grid[7], int threads);
/* The 3D sort key: the stencil base cell (Flip3.Stencil3.of) renumbered brick-major,
 * 8x8x8-cell bricks; a slot with m <= 0 takes orc_mpm3_inactive_key, one past every cell's key.
 * orc_mpm3_key_bits returns the number of key bits. The reference's 3D step has no sort: this
 * is the key this build sorts by. */
int32_t orc_mpm3_inactive_key(int nx, int ny, int nz);
int orc_mpm3_key_bits(int nx, int ny, int nz);
void orc_mpm3_keys(int nx, int ny, int nz, int64_t count, const float* x, const float* y,
                   const float* z, const float* m, int32_t* keys);

/* 3D surface tension: Tension3.blur (core/particle/Tension3.java:69-107; scaled: the mass to a
 * colour), Tension3.stress (:122-147; t = txx tyy tzz txy txz tyz), Tension3.force (:156-196)
 * at one node, and Flip3.gridWithTension (core/particle/Flip3.java:263-311). */
void orc_tension3_blur(int nx, int ny, int nz, const float* src, float* dst, const int32_t params[7], int scaled);
void orc_tension3_stress(int nx, int ny, int nz, const float* colour, float* const t[6]);
void orc_tension3_force(int nx, int ny, int nz, const float* colour, const float* const t[6], int64_t node,
                        float out[3]);
void orc_mpm3_grid_tension(int nx, int ny, int nz, const int32_t params[7], const float* gm, const float* gmu,
                           const float* gmv, const float* gmw, float* gu, float* gv, float* gw,
                           const float* const t[6], const float* colour);
/* 2D surface tension: Tension.blur (core/particle/Tension.java:98-122; scaled: min(m / rho0, 1)
 * per tap), Tension.stress (:135-150), Tension.force (:160-185) at one node, and
 * Flip.gridWithTension (core/particle/Flip.java:342-412); params are Flip's 28 words. */
void orc_tension2_blur(int nx, int ny, const float* src, float* dst, const int32_t* params, int scaled);
void orc_tension2_stress(int nx, int ny, const float* colour, float* txx, float* tyy, float* txy);
void orc_tension2_force(int nx, int ny, const float* colour, const float* txx, const float* tyy, const float* txy,
                        int64_t node, float out[2]);
void orc_mpm2_grid_tension(int nx, int ny, const int32_t* params, const float* gm, const float* gmu, const float* gmv,
                           float* gu, float* gv, const float* txx, const float* tyy, const float* txy,
                           const float* colour);

/* The 3D level of detail: Flip3.lod (core/particle/Flip3.java:420-586) over `groups` groups of 8
 * consecutive slots, with level (per group), target (per group), stay (per slot); and
 * Flip3.lodTarget (:600-641) with view = eye x, y, z, near, floor. */
void orc_mpm3_lod(int nx, int ny, int nz, int64_t groups, float* const p[17], float* level, const float* target,
                  float* stay);
void orc_mpm3_lod_target(int64_t groups, const float* x, const float* y, const float* z, const float* level,
                         float* target, const float view[5]);

/* ---- 2D MLS-MPM / APIC, core/particle/Flip.java ----
 * Particle fields in FlipStep order: x y u v m j c00 c01 c10 c11 (10). Grid: gm gmu gmv, gu gv. */
void orc_mpm2_clear(int nx, int ny, float* gm, float* gmu, float* gmv);
void orc_mpm2_scatter(int nx, int ny, int64_t count, float* const p[10], const int32_t* params,
                      float* gm, float* gmu, float* gmv);
void orc_mpm2_grid(int nx, int ny, const int32_t* params, const float* gm, const float* gmu,
                   const float* gmv, float* gu, float* gv);
void orc_mpm2_advect(int nx, int ny, int64_t count, float* const p[10], const int32_t* params,
                     const float* gu, const float* gv);
void orc_mpm2_step(int nx, int ny, int64_t count, float* const p[10], const int32_t* params,
                   float* const grid[5]);
/* The deterministic P2G: each of a particle's 27 amounts (computed as orc_mpm2_scatter computes
 * them) is rounded to a fixed-point integer -- mass in units of 2^-40, momentum in 2^-32, round to
 * nearest even -- and summed in 64-bit integers, which are associative, so every order and
 * grouping of the sums gives the same grid (the principle of core/particle/FixedScatter.java:24-31,
 * applied to the quadratic APIC amounts). orc_mpm2_decode_fixed turns the sums back into the
 * float accumulators the grid update reads: (float)((double)q * 2^-bits). */
#define ORC_FIXED_MASS_BITS 40
#define ORC_FIXED_MOMENTUM_BITS 32
void orc_mpm2_scatter_fixed(int nx, int ny, int64_t count, float* const p[10], const int32_t* params,
                            int64_t* qm, int64_t* qmu, int64_t* qmv);
void orc_mpm2_decode_fixed(int64_t nodes, const int64_t* qm, const int64_t* qmu, const int64_t* qmv,
                           float* gm, float* gmu, float* gmv);
/* core/particle/Scatter.java:473-477 */
int32_t orc_mpm2_cell(float x, float y, int nx, int ny);
void orc_mpm2_keys(int nx, int ny, int64_t count, const float* x, const float* y, int32_t* keys);
/* core/particle/Scatter.java:484-502; returns 0, or -1-p for the first particle out of order. */
int64_t orc_mpm2_cell_starts(int nx, int ny, int64_t count, const float* x, const float* y,
                             int32_t* starts);
/* The sort by cell with the copy back (core/particle/FlipStep.java:183-192), made stable. */
void orc_mpm2_sort(int nx, int ny, int64_t count, float* const p[10]);

/* ParticleDiagnostics.of (core/particle/ParticleDiagnostics.java:45-65): out = min density,
 * max density, low (99% above), high (99% below), mach, courant, non-finite count. */
void orc_particle_diagnostics(int64_t count, const float* j, double speed, double sound, double dt, double out[7]);

/* ---- the fixed-point bilinear scatter (core/particle/FixedScatter.java) ----
 * orc_fixed_scatter adds into i32 grids (wrapping) what FixedScatter.kernel/segmented deposit:
 * test/particle/FixedScatterTest.java:402-432. orc_fixed_split is Quanta's share,
 * floor(q * M / 2^20) in i32 (FixedScatter.java:186-193). orc_scatter_f64 is the FP64 bilinear
 * deposit of test/particle/ScatterTest.java:429-450. */
#define ORC_FIXED_OFFSET_BITS 10
#define ORC_FIXED_MAX_MASS (1 << 21)
int32_t orc_fixed_split(int32_t q, int32_t mass);
void orc_fixed_scatter(int n, int64_t count, const float* x, const float* y, const float* u, const float* v,
                       const float* m, int mass_bits, int velocity_bits, int32_t* gm, int32_t* gmu, int32_t* gmv);
void orc_scatter_f64(int n, int64_t count, const float* x, const float* y, const float* u, const float* v,
                     const float* m, double* gm, double* gmu, double* gmv);

/* ---- stable sort by key: order[i] = index of the i-th particle in key order ---- */
void orc_stable_order(int64_t count, const int32_t* keys, int32_t* order);
/* Exclusive starts of `cells` buckets, cells + 1 words (the last is `count`). */
void orc_bucket_starts(int64_t count, const int32_t* keys, int32_t cells, int32_t* starts);

#ifdef __cplusplus
}
#endif
#endif
