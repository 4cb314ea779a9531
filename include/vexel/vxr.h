/*
 * vxr.h -- the C ABI of libvexelhip.so: the MI355X-native hot path of vexelray-sim-fluid
 * (the HLL shallow-water step and the MLS-MPM/APIC substep) behind the reference's
 * simulator API, for a Java FFM (or any C FFI) host.
 *
 * Every entry point returns a status code (VXR_OK or an error below) and leaves a message
 * for vxr_last_error() on failure. Plain pointers, sizes and 32-bit words only.
 *
 * Reference paths are shortened: core/ = vexelray-sim-fluid-core/src/main/java/dev/vexelray/sim/fluid/,
 * gui/ = vexelray-sim-fluid-gui/src/main/java/dev/vexelray/sim/fluid/gui/.
 */
#ifndef VEXEL_VXR_H
#define VEXEL_VXR_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. The Java shim maps EINVAL to IllegalArgumentException and ESTATE to
 * IllegalStateException, as the reference throws them (core/stencil/ShallowWater.java:167-169,
 * core/particle/FlipStep.java:83-88, gui/ParticleSimulation.java:147-149). */
#define VXR_OK 0
#define VXR_EINVAL 1       /* bad shape, name, length or parameter */
#define VXR_ESTATE 2       /* a call that does not fit the simulation's kind or state */
#define VXR_EDEVICE 3      /* a HIP error; the message carries hipGetErrorString */
#define VXR_ENOMEM 4       /* device allocation failed */
#define VXR_EUNSUPPORTED 5 /* a flag outside the built path (tension, heat, pump, ...) */

/* Simulation kinds. */
#define VXR_KIND_MPM2 1 /* gui/ParticleSimulation.java over core/particle/FlipStep.java */
#define VXR_KIND_MPM3 2 /* gui/ParticleSimulation3.java over core/particle/Flip3Step.java */
#define VXR_KIND_SWE 3  /* gui/PatchSimulation.java over core/stencil/ShallowWater.java */
#define VXR_KIND_MPM2_BATCH 4 /* many ParticleSimulation scenes in one set of buffers (vxr_create_batch) */

/* Creation flags. Of the reference's optional physics, surface tension is built. MPM3 with
 * VXR_FLAG_TENSION steps as Flip3Step(..., tension) does (core/particle/Flip3Step.java:88-109:
 * clear, scatter, blur0..blur5, stress, grid with the capillary force, advect; sigma is params[6];
 * buffers c0 c1 txx tyy tzz txy txz tyz); MPM2 with it as FlipStep(..., tension) does
 * (core/particle/FlipStep.java:111-143: the same passes; sigma is params[5]; buffers c0 c1 txx
 * tyy txy). FlipStep's heat, convection, relax, pump and Flip3Step's plastic are
 * VXR_EUNSUPPORTED. */
#define VXR_FLAG_TENSION (1u << 0)
#define VXR_FLAG_HEAT (1u << 1)
#define VXR_FLAG_CONVECTION (1u << 2)
#define VXR_FLAG_RELAX (1u << 3)
#define VXR_FLAG_PUMP (1u << 4)
#define VXR_FLAG_PLASTIC (1u << 5)
/* Shallow-water edges (core/stencil/Edges.java): set = OPEN, clear = WALL. */
#define VXR_FLAG_OPEN_WEST (1u << 16)
#define VXR_FLAG_OPEN_EAST (1u << 17)
#define VXR_FLAG_OPEN_SOUTH (1u << 18)
#define VXR_FLAG_OPEN_NORTH (1u << 19)

/* Pass groups for vxr_run: the step (clear, scatter, grid, advect) and the sort by cell
 * (keys, radix passes, permutation of every particle field). */
#define VXR_GROUP_STEP 0
#define VXR_GROUP_SORT 1
#define VXR_GROUP_LOD 2 /* MPM3: lod_target then lod (gui/ParticleSimulation3.java:134-135,346-349) */

typedef struct vxr_sim vxr_sim;

/* The message of the last failed call on this thread; never NULL. */
const char* vxr_last_error(void);

/* Number of HIP devices visible; *n = 0 and VXR_OK when there is none. */
int vxr_device_count(int* n);

/* Makes a simulation on the current HIP device with every buffer allocated and zeroed
 * (PassRunner.gpu + runner.clear(), gui/ParticleSimulation.java:76-80):
 *   MPM2: dims = {nx, ny, -}, nx, ny >= 4 (core/particle/FlipStep.java:81-193)
 *   MPM3: dims = {nx, ny, nz}, each >= 4 (core/particle/Flip3Step.java:53-123)
 *   SWE:  dims = {nx, ny, -}, each >= 1, particles ignored (gui/PatchSimulation.java:58-76)
 * Safe on any thread; the handle then belongs to the thread that uses it. */
int vxr_create(int kind, const int32_t dims[3], int64_t particles, uint32_t flags, vxr_sim** out);

/* Makes a batch of `scenes` independent 2D scenes (no reference counterpart: the reference runs
 * one gui/ParticleSimulation per scene). Scene k has an nx[k] x ny[k] grid and particles[k]
 * particles. Every pass runs all scenes in one launch, and each scene's results are those of a
 * single-scene VXR_KIND_MPM2 run with option p2g = 2 (the fixed-point P2G), bit for bit. The
 * buffers have the 2D names; each holds every scene at the offsets vxr_batch_layout gives
 * (particle ranges padded with m = 0 particles), and "params" holds 28 words per scene. */
int vxr_create_batch(int scenes, const int32_t* nx, const int32_t* ny, const int64_t* particles, uint32_t flags,
                     vxr_sim** out);

/* Scene k's first particle, node and start in the batch's buffers; each array takes scenes + 1
 * entries, the last being the buffer's length. */
int vxr_batch_layout(vxr_sim* sim, int64_t* particle_offsets, int64_t* node_offsets, int64_t* start_offsets);

/* Scenes in a batch (1 for any other simulation). */
int vxr_scene_count(vxr_sim* sim, int* n);

/* Frees every device buffer and the stream (ParticleSimulation.close, :228-231). */
int vxr_destroy(vxr_sim* sim);

/* Writes `n` 32-bit words into the named buffer, synchronously, from host memory the call
 * only borrows (PassRunner.write / ResidentBuffer.write). `n` must equal the buffer's
 * length; "params" also takes a shorter prefix. Names: MPM3 x y z u v w m j c00..c22 gm gmu
 * gmv gmw gu gv gw params keys level target stay view slot active (and with VXR_FLAG_TENSION
 * c0 c1 txx tyy tzz txy txz tyz); MPM2 x y u v m j c00 c01 c10 c11 gm gmu gmv gu gv params keys
 * starts (and with VXR_FLAG_TENSION c0 c1 txx tyy txy); SWE h hu hv params speed budget clamped (the current slot of each rotating set,
 * as gui/PatchSimulation.java:147-213 resolves it). MPM3 "slot" (int32) is the slot each
 * particle position holds -- the identity until a sort permutes it with the particles -- and a
 * write of it must be a permutation of 0..count-1; "active" (int32, read-only) is the number of
 * leading particles the last sort found with mass (count when unknown). */
int vxr_write(vxr_sim* sim, const char* name, const void* words, int64_t n);

/* Reads `n` words of the named buffer after every queued pass has finished (a readback,
 * as PassRunner.read / ResidentBuffer.read). */
int vxr_read(vxr_sim* sim, const char* name, void* words, int64_t n);

/* The named buffer's length in words. */
int vxr_length(vxr_sim* sim, const char* name, int64_t* n);

/* Queues `repeat` runs of a pass group (PassRunner.run(step.step()) / run(step.sort()),
 * gui/ParticleSimulation.java:175-183, gui/ParticleSimulation3.java:291-296). Asynchronous. */
int vxr_run(vxr_sim* sim, int group, int repeat);

/* ParticleSimulation.advance (gui/ParticleSimulation.java:175-183): `count` steps, numbered from
 * `first_step` (the caller's step counter), with a sort first on each step whose number is a
 * multiple of `sort_every` (0: never). With option "graph" = 1 and timing off, each run of a sort
 * and the steps up to the next one is captured once as a hipGraph and replayed. Asynchronous. */
int vxr_advance(vxr_sim* sim, int count, int sort_every, int64_t first_step);

/* Queues one named pass: "clear", "scatter", "grid", "advect", "sort" (a sub-list of the
 * step, as test/particle/Flip3Test.java:102 runs step().subList(0, 2)); MPM3 also "lod_target" and
 * "lod" (core/particle/Flip3.java:420-641) over count / 8 groups of 8 consecutive slots, with the
 * buffers "level", "target" (a word a group), "stay" (a word a slot) and "view" (eye x y z, near,
 * floor). The groups are the slots as loaded; after a sort the passes find a slot's particle
 * through the slot map, and the sort moves the slots without mass behind every active one, where
 * P2G and G2P no longer visit them (until a lod pass or a write of "m" may give them mass). */
int vxr_run_pass(vxr_sim* sim, const char* pass, int repeat);

/* SWE only: queues `count` steps with the reference's buffer rotation -- state k%2 -> (k+1)%2,
 * speed read k%3 / max (k+1)%3 / clear (k+2)%3, budget k%2 -> (k+1)%2
 * (gui/PatchSimulation.java:197-205, core/stencil/ShallowWater.java:92-94). */
int vxr_swe_dispatch(vxr_sim* sim, int count);

/* Diagnostics.of (core/stencil/Diagnostics.java:48-81) of the current shallow-water state, on the
 * device: out = volume, min depth, max depth, max speed, max Froude, max Courant, non-finite,
 * negative and wet cell counts. Minima, maxima and counts are the sequential loop's exactly; the
 * volume is summed in a fixed tree order. Synchronous. */
int vxr_swe_diagnostics(vxr_sim* sim, double g, double dry, double dt, double dx, double dy, double out[9]);

/* ParticleDiagnostics.of (core/particle/ParticleDiagnostics.java:45-65) of the particles' J, on
 * the device: out = min density, max density, low density (99% above), high density (99% below),
 * mach, courant, non-finite count. Reuses the sort's key buffers ("keys" then holds J's bits).
 * For a batch, over every scene's particles together. Synchronous. */
int vxr_particle_diagnostics(vxr_sim* sim, double speed, double sound, double dt, double out[7]);

/* vxr_particle_diagnostics without allocation. No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc). ParticleDiagnostics.of of `count` device floats `j` (J), queued on
 * `stream` (a hipStream_t; 0 the null stream) and read back synchronously into out[7] as
 * vxr_particle_diagnostics gives it. Its sort lives in a fixed device array of the library
 * (16.3 MB, held by each GPU it runs on until the process ends); count must be 1 to 1,048,576
 * (VXR_EINVAL past that). Calls that overlap on one GPU share that array and race. */
int vxr_particle_diagnostics_noalloc(void* stream, const float* j, int64_t count, double speed, double sound,
                                     double dt, double out[7]);

/* The fixed-point bilinear scatter (core/particle/FixedScatter.java:151-183: FixedScatter.kernel,
 * schedule 0, and FixedScatter.segmented, schedule 1) of `count` particles (device arrays x y u v m,
 * node units) into the i32 grids gridMq, gridMuq, gridMvq of an n x n grid (device arrays, added to),
 * at the scale FixedScatter.FixedPoint(mass_bits, velocity_bits) (:91-148; each 0..30). Queued on
 * `stream` (a hipStream_t; 0 the null stream). Every schedule and order gives the same bits. */
int vxr_fixed_scatter(void* stream, int n, int64_t count, const float* x, const float* y, const float* u,
                      const float* v, const float* m, int mass_bits, int velocity_bits, int schedule, int32_t* grid_m,
                      int32_t* grid_mu, int32_t* grid_mv);

/* Zeroes every buffer (PassRunner.clear). */
int vxr_clear(vxr_sim* sim);

/* Blocks until every queued pass has finished (ParticleSimulation3.finish, :400-402). */
int vxr_finish(vxr_sim* sim);

/* Lends the device address of a named buffer (ParticleSimulation3.vkBuffer, :377-383). Grid
 * and SWE addresses are valid until vxr_destroy; particle fields move to a new buffer when
 * a sort permutes them, so a particle address is valid until the next sort. For the SWE
 * state the address is that of the current slot, which changes with each dispatch. */
int vxr_device_ptr(vxr_sim* sim, const char* name, void** ptr);

/* The hipStream_t the simulation queues on, for a caller that orders its own work after it. */
int vxr_stream(vxr_sim* sim, void** stream);

/* Schedule knobs (no reference counterpart; vxr_autotune sets them):
 *   "p2g"        MPM: 0 = direct float atomics (the reference's Flip3.scatter schedule),
 *                1 = sorted-segmented (default): per-thread runs of consecutive particles folded in
 *                registers (3D: then a wave-private LDS node tile), one atomic per node and field;
 *                2 = fixed point (2D, 3D, and always for a batch): each amount rounded to a 64-bit
 *                integer (mass 2^-40, momentum 2^-32) and summed by integer atomics -- the same
 *                bits in any order, equal to the CPU restatement's, so that with g2p 0 or 1 whole
 *                substeps are deterministic (3D: the sums take 32 bytes a node, made when the
 *                option is first set);
 *                3 = MPM3 only: cell-ranked groups -- the sort also lays each 256-cell group out by
 *                cells ranked by particle count, a wave takes a group with a lane a cell, folds
 *                each cell's particles in registers and adds them into an LDS node tile in 27
 *                conflict-free steps; a particle that changed cell since the sort goes by direct
 *                atomics, and until the first sort (or after a write of positions, masses, slots
 *                or keys, or a lod pass) the scatter runs as option 1
 *   "g2p"        MPM3: 0 = per-particle gathers and 1 = an LDS node tile, both in the reference's
 *                arithmetic order (the CPU restatement's bits); 2 = the LDS tile with the sums
 *                contracted axis by axis in fused multiply-adds (default, within a few ulps)
 *   "run"        MPM: consecutive particles a thread takes in the segmented passes: 1, 2 or 4
 *                (default 4); for the 2D fixed-point scatter, the particles folded while they
 *                share a cell in blocks whose amounts reach 2^40 (other blocks add each amount)
 *   "p2g_ints"   MPM3 with p2g 1: 1 = each wave's LDS node tile sums in 32-bit integers, in
 *                power-of-two units the wave chooses from bounds on its particles' amounts (mass
 *                and momentum apart), each run of particles rounded once to the nearest unit and
 *                added by LDS integer atomics; 0 = float tiles added by read-add-write rounds.
 *                Both within FP32 reassociation of the reference's sums (default 1)
 *   "radix_bits" sort digit width (4..8, default 8)
 *   "rows"       SWE: rows a wave marches per tile (1..256, default 32)
 *   "hll_fast"   SWE: 1 = the step with one refined hardware reciprocal per face and per cell in
 *                place of its correctly rounded divisions, and the hardware square root: fewer
 *                instructions, within a few ulps a step of the reference's bits (default 0: the
 *                reference's bits)
 *   "blocks"     2D fixed-point and batch passes: blocks per launch, each striding over its
 *                items (particle chunks, or windows of node flags for clear and grid; 0 = one
 *                block per item)
 *   "graph"      1 = vxr_advance replays captured hipGraphs (default 0) */
int vxr_set_option(vxr_sim* sim, const char* key, int64_t value);
int vxr_get_option(vxr_sim* sim, const char* key, int64_t* value);

/* Times the choices of "run", "blocks" and "radix_bits" (MPM) or "rows" (SWE) on this
 * simulation's own data, `reps` runs each, and keeps the fastest. Only passes that leave the
 * state as it was are timed: clear, scatter and grid, the sort's keys and radix passes without
 * the permutation, and the shallow-water step into the spare slot. Synchronous. */
int vxr_autotune(vxr_sim* sim, int reps);

/* Kernel timing with HIP events on the simulation's own stream. Enabling clears the
 * totals; vxr_kernel_time synchronizes and returns the summed milliseconds and launch
 * count of every launch of the named kernel since then. */
int vxr_timing(vxr_sim* sim, int enable);
/* Timing of the named kernel alone (an event pair around each of its launches and none around the
 * others, whose launches then follow each other with no event between them); clears the totals.
 * vxr_timing(sim, 0) turns it off. */
int vxr_timing_only(vxr_sim* sim, const char* kernel);
int vxr_kernel_time(vxr_sim* sim, const char* kernel, double* total_ms, int64_t* launches);

#ifdef __cplusplus
}
#endif
#endif
