package com.cartogenesis.desktop

import com.cartogenesis.worldgen.pipeline.OceanAccelerator
import com.cartogenesis.worldgen.pipeline.OceanCirculation
import com.cartogenesis.worldgen.pipeline.OceanStencil
import java.nio.ByteBuffer
import java.nio.ByteOrder
import kotlinx.coroutines.Job
import kotlinx.coroutines.currentCoroutineContext

/**
 * The MI355X ocean solver behind the reference's unchanged seam, [OceanAccelerator]
 * (worldgen/.../pipeline/OceanAccelerator.kt:15-42), over libcgocean.so's C ABI (include/cgocean.h)
 * through the JNI shim cgocean_jni.c beside this file.
 *
 * Not compiled in this repository (it has no JVM); a maintainer drops it into desktop/src/main/kotlin
 * in place of GpuOcean.kt and registers it where DesktopPlatform.kt:180 creates GpuOcean:
 * `oceanProbe = HipOcean.createOrNull()`.
 *
 * Two capabilities:
 *  - [solve], the seam: batches of red-black passes, bit-identical to OceanCirculation.relax
 *    (OceanCirculation.kt:241-274); null means "relax on the processor", as for GpuOcean;
 *  - [solveLevels] and [Round], the whole solve on the device, which the seam cannot express and
 *    which removes the per-batch host round trip (docs/TODO.md:2119-2126). circulateUnder
 *    (OceanStage.kt:417-498) asks for it where present and falls back to the processor's solve
 *    when it returns null.
 *
 * `OceanStencil.centreWeight` is internal to worldgen; the whole solve needs it, so the maintainer
 * widens it to public (the seam alone does not: the upload passes null then).
 */
class HipOcean private constructor(private val ctx: Long, override val name: String) : OceanAccelerator {

    class Result(val accelerator: HipOcean?, val unavailableBecause: String?)

    /** Resident stencils, keyed on the identity of their east weights, as GpuOcean.holding keys them (GpuOcean.kt:104-137). */
    private val resident = ArrayDeque<Pair<FloatArray, Long>>()

    override suspend fun solve(stencil: OceanStencil, start: FloatArray, passes: Int): FloatArray? {
        val across = stencil.cellsAcross
        if (across <= 0 || across % 2 != 0 || stencil.cellsDown <= 0 || passes < 0) return null
        if (start.size != across * stencil.cellsDown) return null
        return withCancel { cancel ->
            synchronized(this) {
                val h = holding(stencil, withCentre = false) ?: return@synchronized null
                val out = FloatArray(start.size)
                if (nRelax(ctx, h, stencil.forcing, start, out, passes, cancel) == CG_OK) out else null
            }
        }
    }

    /**
     * OceanCirculation.solve / solveWithKrylovRescue (OceanCirculation.kt:473-528) on the device over
     * [levels], finest first; [method] is one of the CG_METHOD_* codes. Null when the device declines
     * or fails: the caller solves on the processor. requireSolved stays with the caller.
     */
    suspend fun solveLevels(
        levels: List<OceanStencil>,
        start: FloatArray?,
        tolerance: Double,
        poleIsWall: Boolean,
        method: Int = CG_METHOD_REFERENCE
    ): OceanCirculation.Solution? = withCancel { cancel ->
        synchronized(this) {
            val handles = LongArray(levels.size) { holding(levels[it], withCentre = true) ?: return@synchronized null }
            val out = FloatArray(levels.first().forcing.size)
            val info = DoubleArray(3)
            val rc = nSolve(ctx, handles, levels.first().forcing, start, tolerance, poleIsWall, method, out, info, cancel)
            if (rc != CG_OK) null else OceanCirculation.Solution(out, info[0].toInt(), info[1])
        }
    }

    /**
     * A resident round (OceanStage.circulateUnder, OceanStage.kt:417-498): one plan for the
     * circulation's ladder and one for the heat's, reused across rounds on the same coast. Each
     * round uploads only the stress (curl and withBalance on the device) and the heat's start, and
     * reads back the stream, the velocities and the temperature.
     */
    inner class Round(flowLevels: List<OceanStencil>) : AutoCloseable {
        private val flow = nPlanCreate(ctx, LongArray(flowLevels.size) { holding(flowLevels[it], true) ?: 0L }, true)
            .also { require(it != 0L) { "the device could not hold the circulation's grids: ${nLastError(ctx)}" } }
        private val cells = flowLevels.first().forcing.size

        /** The stream function under a stress, [method] as in [solveLevels]; null to fall back. */
        fun circulation(
            widthMeters: Double, heightMeters: Double, eastward: DoubleArray, northward: DoubleArray,
            tolerance: Double, method: Int = CG_METHOD_REFERENCE
        ): OceanCirculation.Solution? = synchronized(this@HipOcean) {
            if (nPlanSetStress(flow, widthMeters, heightMeters, eastward, northward, null) != CG_OK) return null
            val info = DoubleArray(3)
            if (nPlanSolve(flow, tolerance, method, false, info) != CG_OK) return null
            val stream = FloatArray(cells)
            if (nPlanRead(flow, stream) != CG_OK) return null
            OceanCirculation.Solution(stream, info[0].toInt(), info[1])
        }

        /**
         * The learned coarse-grid correction for this round's circulation: a trained U-Net's bf16
         * weights ([weights], as cg_unet_weight_count(base) shorts, layer by layer) on the ladder's grid
         * [level], with that grid's depth in metres; afterwards `circulation(..., method =
         * CG_METHOD_LEARNED)` runs it inside each refinement step. Null weights remove it.
         */
        fun setCoarseNet(level: Int, base: Int, weights: ShortArray?, depth: FloatArray?, outScale: Double): Boolean =
            synchronized(this@HipOcean) { nPlanSetCoarseNet(flow, level, base, weights, depth, outScale) == CG_OK }

        /** OceanCirculation.velocities (OceanCirculation.kt:801-828) of the last stream. */
        fun velocities(widthMeters: Double, heightMeters: Double, eastward: FloatArray, northward: FloatArray): Boolean =
            synchronized(this@HipOcean) { nPlanVelocities(flow, widthMeters, heightMeters, eastward, northward) == CG_OK }

        /**
         * The heat (OceanStage.kt:486-497): solveByKrylov with [bodies], restarted up to three times,
         * from [start], pole not a wall.
         */
        fun heat(
            heatLevels: List<OceanStencil>, bodies: OceanCirculation.WaterBodies, start: FloatArray, tolerance: Double
        ): OceanCirculation.Solution? = synchronized(this@HipOcean) {
            val handles = LongArray(heatLevels.size) { holding(heatLevels[it], true) ?: return null }
            val plan = nPlanCreate(ctx, handles, false)
            if (plan == 0L) return null
            try {
                if (nPlanSetBodies(plan, bodies.bodyOf, bodies.count) != CG_OK) return null
                if (nPlanSetStart(plan, start) != CG_OK) return null
                val info = DoubleArray(3)
                if (nPlanSolve(plan, tolerance, CG_METHOD_KRYLOV, true, info) != CG_OK) return null
                val values = FloatArray(start.size)
                if (nPlanRead(plan, values) != CG_OK) return null
                OceanCirculation.Solution(values, info[0].toInt(), info[1])
            } finally {
                nPlanDestroy(plan)
            }
        }

        override fun close() = synchronized(this@HipOcean) { nPlanDestroy(flow) }
    }

    private fun holding(stencil: OceanStencil, withCentre: Boolean): Long? {
        resident.firstOrNull { it.first === stencil.eastWeight }?.let { entry ->
            resident.remove(entry)
            resident.addLast(entry)
            return entry.second
        }
        while (resident.size >= MOST_RESIDENT_STENCILS) nRelease(ctx, resident.removeFirst().second)
        val water = ByteArray(stencil.isWater.size) { if (stencil.isWater[it]) 1 else 0 }
        val h = nUpload(
            ctx, stencil.cellsAcross, stencil.cellsDown, water, stencil.eastWeight, stencil.westWeight,
            stencil.northWeight, stencil.southWeight, if (withCentre) stencil.centreWeight else null
        )
        if (h == 0L) return null
        resident.addLast(stencil.eastWeight to h)
        return h
    }

    /** Runs [block] with a direct one-int buffer the library polls, set when the coroutine is cancelled (GpuOcean.kt:65, 85-86). */
    private suspend inline fun <T> withCancel(block: (ByteBuffer) -> T): T {
        val cancel = ByteBuffer.allocateDirect(4).order(ByteOrder.nativeOrder())
        val job = currentCoroutineContext()[Job]
        val handle = job?.invokeOnCompletion { if (job.isCancelled) cancel.putInt(0, 1) }
        try {
            return block(cancel)
        } finally {
            handle?.dispose()
        }
    }

    companion object {
        const val CG_OK = 0
        const val CG_METHOD_REFERENCE = 0
        const val CG_METHOD_RESCUE = 1
        const val CG_METHOD_REFINED = 2
        const val CG_METHOD_KRYLOV = 4
        const val CG_METHOD_LEARNED = 5
        private const val MOST_RESIDENT_STENCILS = 24

        /** GpuOcean.createOrNull's contract (GpuOcean.kt:189-202): an accelerator, or why there is none. */
        fun createOrNull(device: Int = 0): Result = try {
            System.loadLibrary("cgocean_jni")
            val reason = arrayOfNulls<String>(1)
            val ctx = nOpen(device, reason)
            if (ctx == 0L) Result(null, reason[0] ?: "no HIP device") else Result(HipOcean(ctx, nName(ctx)), null)
        } catch (e: UnsatisfiedLinkError) {
            Result(null, "the HIP ocean library is not installed: ${e.message}")
        }

        @JvmStatic private external fun nOpen(device: Int, reason: Array<String?>): Long
        @JvmStatic private external fun nClose(ctx: Long)
        @JvmStatic private external fun nName(ctx: Long): String
        @JvmStatic private external fun nLastError(ctx: Long): String
        @JvmStatic private external fun nUpload(
            ctx: Long, across: Int, down: Int, water: ByteArray, e: FloatArray, w: FloatArray, n: FloatArray,
            s: FloatArray, centre: DoubleArray?
        ): Long
        @JvmStatic private external fun nRelease(ctx: Long, handle: Long)
        @JvmStatic private external fun nRelax(
            ctx: Long, handle: Long, forcing: FloatArray, start: FloatArray, out: FloatArray, passes: Int, cancel: ByteBuffer
        ): Int
        @JvmStatic private external fun nSolve(
            ctx: Long, levels: LongArray, forcing: FloatArray, start: FloatArray?, tol: Double, poleIsWall: Boolean,
            method: Int, out: FloatArray, info: DoubleArray, cancel: ByteBuffer
        ): Int
        @JvmStatic private external fun nPlanCreate(ctx: Long, levels: LongArray, poleIsWall: Boolean): Long
        @JvmStatic private external fun nPlanDestroy(plan: Long)
        @JvmStatic private external fun nPlanSetForcing(plan: Long, forcing: FloatArray): Int
        @JvmStatic private external fun nPlanSetStress(
            plan: Long, widthM: Double, heightM: Double, eastward: DoubleArray, northward: DoubleArray, depth: DoubleArray?
        ): Int
        @JvmStatic private external fun nPlanSetBodies(plan: Long, bodyOf: IntArray?, count: Int): Int
        @JvmStatic private external fun nPlanSetStart(plan: Long, start: FloatArray): Int
        @JvmStatic private external fun nPlanSolve(plan: Long, tol: Double, method: Int, warm: Boolean, info: DoubleArray): Int
        @JvmStatic private external fun nPlanRead(plan: Long, out: FloatArray): Int
        @JvmStatic private external fun nPlanVelocities(
            plan: Long, widthM: Double, heightM: Double, eastward: FloatArray, northward: FloatArray
        ): Int
        @JvmStatic private external fun nPlanSetCoarseNet(
            plan: Long, level: Int, base: Int, weights: ShortArray?, depth: FloatArray?, outScale: Double
        ): Int
    }
}
