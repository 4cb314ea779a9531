/*
 * PassRunner's surface over libvexelhip.so for the reference's GUI runners, through the Java FFM
 * API (java.lang.foreign, Java 22+). The C ABI it binds is include/vexel/vxr.h.
 *
 * Source only: this image has no JDK, so the file is not compiled here. A maintainer adds it
 * next to gui/ParticleSimulation3.java and changes the runner classes in their constructors
 * (gui/ParticleSimulation.java:77, gui/ParticleSimulation3.java:133,
 * gui/PatchSimulation.java:58-76), where PassRunner.gpu(accelerator, step, 256, 32) is made today.
 */
package dev.vexelray.sim.fluid.gui;

import static java.lang.foreign.ValueLayout.ADDRESS;
import static java.lang.foreign.ValueLayout.JAVA_BYTE;
import static java.lang.foreign.ValueLayout.JAVA_DOUBLE;
import static java.lang.foreign.ValueLayout.JAVA_FLOAT;
import static java.lang.foreign.ValueLayout.JAVA_INT;
import static java.lang.foreign.ValueLayout.JAVA_LONG;

import java.lang.foreign.Arena;
import java.lang.foreign.FunctionDescriptor;
import java.lang.foreign.Linker;
import java.lang.foreign.MemoryLayout;
import java.lang.foreign.MemorySegment;
import java.lang.foreign.SymbolLookup;
import java.lang.invoke.MethodHandle;
import java.util.List;

/** One simulation on the current HIP device. Owning thread only, as PassRunner is. */
final class NativeRunner implements AutoCloseable {
    /** Simulation kinds (VXR_KIND_*). */
    static final int MPM2 = 1, MPM3 = 2, SWE = 3;
    /** Pass groups (VXR_GROUP_*). */
    static final int STEP = 0, SORT = 1, LOD = 2;
    /** Creation flags (VXR_FLAG_*). */
    static final int TENSION = 1, OPEN_WEST = 1 << 16, OPEN_EAST = 1 << 17, OPEN_SOUTH = 1 << 18, OPEN_NORTH = 1 << 19;

    private static final Linker LINKER = Linker.nativeLinker();
    private static final SymbolLookup LIB = SymbolLookup.libraryLookup("libvexelhip.so", Arena.global());

    private static MethodHandle fn(String name, MemoryLayout... args) {
        return LINKER.downcallHandle(LIB.find(name).orElseThrow(), FunctionDescriptor.of(JAVA_INT, args));
    }

    private static final MethodHandle LAST_ERROR = LINKER.downcallHandle(LIB.find("vxr_last_error").orElseThrow(),
            FunctionDescriptor.of(ADDRESS.withTargetLayout(MemoryLayout.sequenceLayout(4096, JAVA_BYTE))));
    private static final MethodHandle CREATE = fn("vxr_create", JAVA_INT, ADDRESS, JAVA_LONG, JAVA_INT, ADDRESS);
    private static final MethodHandle DESTROY = fn("vxr_destroy", ADDRESS);
    private static final MethodHandle WRITE = fn("vxr_write", ADDRESS, ADDRESS, ADDRESS, JAVA_LONG);
    private static final MethodHandle READ = fn("vxr_read", ADDRESS, ADDRESS, ADDRESS, JAVA_LONG);
    private static final MethodHandle LENGTH = fn("vxr_length", ADDRESS, ADDRESS, ADDRESS);
    private static final MethodHandle RUN = fn("vxr_run", ADDRESS, JAVA_INT, JAVA_INT);
    private static final MethodHandle ADVANCE = fn("vxr_advance", ADDRESS, JAVA_INT, JAVA_INT, JAVA_LONG);
    private static final MethodHandle RUN_PASS = fn("vxr_run_pass", ADDRESS, ADDRESS, JAVA_INT);
    private static final MethodHandle SWE_DISPATCH = fn("vxr_swe_dispatch", ADDRESS, JAVA_INT);
    private static final MethodHandle SWE_DIAGNOSTICS = fn("vxr_swe_diagnostics", ADDRESS, JAVA_DOUBLE, JAVA_DOUBLE,
            JAVA_DOUBLE, JAVA_DOUBLE, JAVA_DOUBLE, ADDRESS);
    private static final MethodHandle PARTICLE_DIAGNOSTICS = fn("vxr_particle_diagnostics", ADDRESS, JAVA_DOUBLE,
            JAVA_DOUBLE, JAVA_DOUBLE, ADDRESS);
    private static final MethodHandle CLEAR = fn("vxr_clear", ADDRESS);
    private static final MethodHandle FINISH = fn("vxr_finish", ADDRESS);
    private static final MethodHandle DEVICE_PTR = fn("vxr_device_ptr", ADDRESS, ADDRESS, ADDRESS);

    private final Arena arena = Arena.ofConfined();
    private final MemorySegment handle;

    /** PassRunner.gpu(accelerator, step, 256, 32): every buffer allocated and zeroed. */
    NativeRunner(int kind, int nx, int ny, int nz, long particles, int flags) {
        MemorySegment dims = arena.allocateFrom(JAVA_INT, nx, ny, nz);
        MemorySegment out = arena.allocate(ADDRESS);
        check(call(CREATE, kind, dims, particles, flags, out));
        handle = out.get(ADDRESS, 0);
    }

    /** runner.write(name, int[]): the whole buffer ("params" also takes a prefix). */
    void write(String name, int[] words) {
        try (Arena a = Arena.ofConfined()) {
            check(call(WRITE, handle, a.allocateFrom(name), a.allocateFrom(JAVA_INT, words), (long) words.length));
        }
    }

    /** runner.write(name, float[]). */
    void write(String name, float[] values) {
        try (Arena a = Arena.ofConfined()) {
            check(call(WRITE, handle, a.allocateFrom(name), a.allocateFrom(JAVA_FLOAT, values), (long) values.length));
        }
    }

    /** runner.floats(name): a readback after every queued pass. */
    float[] floats(String name) {
        try (Arena a = Arena.ofConfined()) {
            MemorySegment cname = a.allocateFrom(name);
            long n = length(a, cname);
            MemorySegment buf = a.allocate(JAVA_FLOAT, n);
            check(call(READ, handle, cname, buf, n));
            return buf.toArray(JAVA_FLOAT);
        }
    }

    /** runner.ints(name). */
    int[] ints(String name) {
        try (Arena a = Arena.ofConfined()) {
            MemorySegment cname = a.allocateFrom(name);
            long n = length(a, cname);
            MemorySegment buf = a.allocate(JAVA_INT, n);
            check(call(READ, handle, cname, buf, n));
            return buf.toArray(JAVA_INT);
        }
    }

    private long length(Arena a, MemorySegment cname) {
        MemorySegment n = a.allocate(JAVA_LONG);
        check(call(LENGTH, handle, cname, n));
        return n.get(JAVA_LONG, 0);
    }

    /** runner.run(passes) for any pass list, name by name (step().subList(0, 2) and the like). */
    void run(List<String> passes) {
        try (Arena a = Arena.ofConfined()) {
            for (String p : passes) check(call(RUN_PASS, handle, a.allocateFrom(p), 1));
        }
    }

    /** runner.run(step.step()), `repeat` times. Asynchronous. */
    void step(int repeat) { check(call(RUN, handle, STEP, repeat)); }

    /** runner.run(step.sort()). Asynchronous. */
    void sort() { check(call(RUN, handle, SORT, 1)); }

    /** ParticleSimulation3's level-of-detail passes (lod_target then lod). */
    void lod() { check(call(RUN, handle, LOD, 1)); }

    /** ParticleSimulation.advance: `count` steps from step number `first`, sorting every `sortEvery`. */
    void advance(int count, int sortEvery, long first) { check(call(ADVANCE, handle, count, sortEvery, first)); }

    /** PatchSimulation.dispatch: `count` HLL steps with the 2/3/2 buffer rotation inside. */
    void dispatch(int count) { check(call(SWE_DISPATCH, handle, count)); }

    /** Diagnostics.of on the device: volume, min/max depth, max speed, Froude, Courant, non-finite, negative, wet. */
    double[] sweDiagnostics(double g, double dry, double dt, double dx, double dy) {
        try (Arena a = Arena.ofConfined()) {
            MemorySegment out = a.allocate(JAVA_DOUBLE, 9);
            check(call(SWE_DIAGNOSTICS, handle, g, dry, dt, dx, dy, out));
            return out.toArray(JAVA_DOUBLE);
        }
    }

    /** ParticleDiagnostics.of on the device: min, max, low, high density, mach, courant, non-finite. */
    double[] particleDiagnostics(double speed, double sound, double dt) {
        try (Arena a = Arena.ofConfined()) {
            MemorySegment out = a.allocate(JAVA_DOUBLE, 7);
            check(call(PARTICLE_DIAGNOSTICS, handle, speed, sound, dt, out));
            return out.toArray(JAVA_DOUBLE);
        }
    }

    /** runner.clear(). */
    void clear() { check(call(CLEAR, handle)); }

    /** runner.finish(): blocks until every queued pass has finished. */
    void finish() { check(call(FINISH, handle)); }

    /** runner.onDevice(): the library has no CPU backend. */
    boolean onDevice() { return true; }

    /** runner.resident(name): a HIP device address (not a Vulkan buffer); particle fields move at each sort. */
    long devicePointer(String name) {
        try (Arena a = Arena.ofConfined()) {
            MemorySegment out = a.allocate(ADDRESS);
            check(call(DEVICE_PTR, handle, a.allocateFrom(name), out));
            return out.get(ADDRESS, 0).address();
        }
    }

    @Override
    public void close() {
        check(call(DESTROY, handle));
        arena.close();
    }

    private static int call(MethodHandle h, Object... args) {
        try {
            return (int) h.invokeWithArguments(args);
        } catch (Throwable t) {
            throw new IllegalStateException(t);
        }
    }

    /** The status codes as the reference throws them: EINVAL, EUNSUPPORTED, the rest IllegalState. */
    private static void check(int status) {
        if (status == 0) return;
        String msg;
        try {
            msg = ((MemorySegment) LAST_ERROR.invoke()).getString(0);
        } catch (Throwable t) {
            msg = "vxr error " + status;
        }
        switch (status) {
            case 1 -> throw new IllegalArgumentException(msg);
            case 5 -> throw new UnsupportedOperationException(msg);
            default -> throw new IllegalStateException(msg);
        }
    }
}
