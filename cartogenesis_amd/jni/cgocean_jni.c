/*
 * cgocean_jni.c: the JNI shim between the Kotlin class com.cartogenesis.desktop.HipOcean
 * (HipOcean.kt beside this file) and the C ABI of libcgocean.so (include/cgocean.h).
 *
 * Not built in this repository: the image has no JDK (no jni.h). A maintainer builds it with
 *   gcc -O2 -shared -fPIC -I$JAVA_HOME/include -I$JAVA_HOME/include/linux -Iinclude \
 *       cartogenesis_amd/jni/cgocean_jni.c -Lcartogenesis_amd -lcgocean -o libcgocean_jni.so
 *
 * Every native copies Java arrays in and out with Get/Release<Type>ArrayElements (no critical
 * regions held across device work), passes plain pointers to the C ABI and returns its status
 * unchanged, so the Kotlin side sees CG_OK / CG_DECLINED / CG_CANCELLED / negative errors exactly as
 * C callers do. A cancel flag is a direct ByteBuffer of one int, written by the coroutine's
 * completion handler and read by the library between launches.
 */
#include <jni.h>
#include <stdint.h>
#include <string.h>

#include "cgocean.h"

#define CTX(h) ((cg_ctx *)(intptr_t)(h))
#define PLAN(h) ((cg_plan *)(intptr_t)(h))

static const volatile int *cancel_flag(JNIEnv *env, jobject buffer) {
    return buffer ? (const volatile int *)(*env)->GetDirectBufferAddress(env, buffer) : NULL;
}

/* Writes {cycles, relative residual, refinements} into a double[3]. */
static void put_info(JNIEnv *env, jdoubleArray info, const cg_solution *s) {
    if (!info) return;
    jdouble v[3] = {(jdouble)s->cycles, s->relative_residual, (jdouble)s->refinements};
    (*env)->SetDoubleArrayRegion(env, info, 0, 3, v);
}

JNIEXPORT jlong JNICALL Java_com_cartogenesis_desktop_HipOcean_nOpen(JNIEnv *env, jclass k, jint device,
                                                                     jobjectArray reason) {
    (void)k;
    cg_ctx *ctx = NULL;
    char why[256] = {0};
    if (cg_ocean_open(device, &ctx, why, sizeof why) != CG_OK) {
        if (reason) (*env)->SetObjectArrayElement(env, reason, 0, (*env)->NewStringUTF(env, why));
        return 0;
    }
    return (jlong)(intptr_t)ctx;
}

JNIEXPORT void JNICALL Java_com_cartogenesis_desktop_HipOcean_nClose(JNIEnv *env, jclass k, jlong ctx) {
    (void)env, (void)k;
    cg_ocean_close(CTX(ctx));
}

JNIEXPORT jstring JNICALL Java_com_cartogenesis_desktop_HipOcean_nName(JNIEnv *env, jclass k, jlong ctx) {
    (void)k;
    char name[256] = {0};
    if (cg_ocean_name(CTX(ctx), name, sizeof name) != CG_OK) strcpy(name, "AMD Instinct GPU");
    return (*env)->NewStringUTF(env, name);
}

JNIEXPORT jstring JNICALL Java_com_cartogenesis_desktop_HipOcean_nLastError(JNIEnv *env, jclass k, jlong ctx) {
    (void)k;
    const char *msg = cg_ocean_last_error(CTX(ctx));
    return (*env)->NewStringUTF(env, msg ? msg : "");
}

/* OceanStencil -> a resident stencil handle (0 on failure). water is 0/1 bytes; centre may be null. */
JNIEXPORT jlong JNICALL Java_com_cartogenesis_desktop_HipOcean_nUpload(JNIEnv *env, jclass k, jlong ctx, jint across,
                                                                       jint down, jbyteArray water, jfloatArray e,
                                                                       jfloatArray w, jfloatArray n, jfloatArray s,
                                                                       jdoubleArray centre) {
    (void)k;
    jbyte *wa = (*env)->GetByteArrayElements(env, water, NULL);
    jfloat *pe = (*env)->GetFloatArrayElements(env, e, NULL);
    jfloat *pw = (*env)->GetFloatArrayElements(env, w, NULL);
    jfloat *pn = (*env)->GetFloatArrayElements(env, n, NULL);
    jfloat *ps = (*env)->GetFloatArrayElements(env, s, NULL);
    jdouble *pc = centre ? (*env)->GetDoubleArrayElements(env, centre, NULL) : NULL;
    cg_stencil handle = 0;
    int rc = cg_ocean_stencil_upload(CTX(ctx), across, down, (const uint8_t *)wa, pe, pw, pn, ps, pc, &handle);
    if (pc) (*env)->ReleaseDoubleArrayElements(env, centre, pc, JNI_ABORT);
    (*env)->ReleaseFloatArrayElements(env, s, ps, JNI_ABORT);
    (*env)->ReleaseFloatArrayElements(env, n, pn, JNI_ABORT);
    (*env)->ReleaseFloatArrayElements(env, w, pw, JNI_ABORT);
    (*env)->ReleaseFloatArrayElements(env, e, pe, JNI_ABORT);
    (*env)->ReleaseByteArrayElements(env, water, wa, JNI_ABORT);
    return rc == CG_OK ? (jlong)handle : 0;
}

JNIEXPORT void JNICALL Java_com_cartogenesis_desktop_HipOcean_nRelease(JNIEnv *env, jclass k, jlong ctx, jlong h) {
    (void)env, (void)k;
    cg_ocean_stencil_release(CTX(ctx), (cg_stencil)h);
}

/* OceanAccelerator.solve: passes of red-black relaxation; start is never written. */
JNIEXPORT jint JNICALL Java_com_cartogenesis_desktop_HipOcean_nRelax(JNIEnv *env, jclass k, jlong ctx, jlong h,
                                                                     jfloatArray forcing, jfloatArray start,
                                                                     jfloatArray out, jint passes, jobject cancel) {
    (void)k;
    jfloat *f = (*env)->GetFloatArrayElements(env, forcing, NULL);
    jfloat *s = (*env)->GetFloatArrayElements(env, start, NULL);
    jfloat *o = (*env)->GetFloatArrayElements(env, out, NULL);
    int rc = cg_ocean_relax(CTX(ctx), (cg_stencil)h, f, s, o, passes, cancel_flag(env, cancel));
    (*env)->ReleaseFloatArrayElements(env, out, o, rc == CG_OK ? 0 : JNI_ABORT);
    (*env)->ReleaseFloatArrayElements(env, start, s, JNI_ABORT);
    (*env)->ReleaseFloatArrayElements(env, forcing, f, JNI_ABORT);
    return rc;
}

/* The whole solve over resident levels (finest first): cg_ocean_solve. start may be null. */
JNIEXPORT jint JNICALL Java_com_cartogenesis_desktop_HipOcean_nSolve(JNIEnv *env, jclass k, jlong ctx,
                                                                     jlongArray levels, jfloatArray forcing,
                                                                     jfloatArray start, jdouble tol,
                                                                     jboolean pole_is_wall, jint method,
                                                                     jfloatArray out, jdoubleArray info,
                                                                     jobject cancel) {
    (void)k;
    const jsize count = (*env)->GetArrayLength(env, levels);
    jlong *hl = (*env)->GetLongArrayElements(env, levels, NULL);
    cg_stencil handles[32];
    if (count < 1 || count > 32) {
        (*env)->ReleaseLongArrayElements(env, levels, hl, JNI_ABORT);
        return CG_ERR_ARG;
    }
    for (jsize i = 0; i < count; i++) handles[i] = (cg_stencil)hl[i];
    (*env)->ReleaseLongArrayElements(env, levels, hl, JNI_ABORT);
    jfloat *f = (*env)->GetFloatArrayElements(env, forcing, NULL);
    jfloat *s = start ? (*env)->GetFloatArrayElements(env, start, NULL) : NULL;
    jfloat *o = (*env)->GetFloatArrayElements(env, out, NULL);
    cg_solution sol = {0, 0.0, 0, 0};
    int rc = cg_ocean_solve(CTX(ctx), handles, count, f, s, tol, pole_is_wall ? 1 : 0, method, o, NULL, &sol,
                            cancel_flag(env, cancel));
    (*env)->ReleaseFloatArrayElements(env, out, o, rc == CG_OK ? 0 : JNI_ABORT);
    if (s) (*env)->ReleaseFloatArrayElements(env, start, s, JNI_ABORT);
    (*env)->ReleaseFloatArrayElements(env, forcing, f, JNI_ABORT);
    put_info(env, info, &sol);
    return rc == CG_OK && sol.status == CG_CANCELLED ? CG_CANCELLED : rc;
}

/* ---- a resident round: one plan per ladder ------------------------------------------------ */

JNIEXPORT jlong JNICALL Java_com_cartogenesis_desktop_HipOcean_nPlanCreate(JNIEnv *env, jclass k, jlong ctx,
                                                                           jlongArray levels,
                                                                           jboolean pole_is_wall) {
    (void)k;
    const jsize count = (*env)->GetArrayLength(env, levels);
    if (count < 1 || count > 32) return 0;
    jlong *hl = (*env)->GetLongArrayElements(env, levels, NULL);
    cg_stencil handles[32];
    for (jsize i = 0; i < count; i++) handles[i] = (cg_stencil)hl[i];
    (*env)->ReleaseLongArrayElements(env, levels, hl, JNI_ABORT);
    cg_plan *plan = NULL;
    if (cg_ocean_plan_create(CTX(ctx), handles, count, pole_is_wall ? 1 : 0, &plan) != CG_OK) return 0;
    return (jlong)(intptr_t)plan;
}

JNIEXPORT void JNICALL Java_com_cartogenesis_desktop_HipOcean_nPlanDestroy(JNIEnv *env, jclass k, jlong plan) {
    (void)env, (void)k;
    cg_ocean_plan_destroy(PLAN(plan));
}

JNIEXPORT jint JNICALL Java_com_cartogenesis_desktop_HipOcean_nPlanSetForcing(JNIEnv *env, jclass k, jlong plan,
                                                                              jfloatArray forcing) {
    (void)k;
    jfloat *f = (*env)->GetFloatArrayElements(env, forcing, NULL);
    int rc = cg_ocean_plan_set_forcing(PLAN(plan), f);
    (*env)->ReleaseFloatArrayElements(env, forcing, f, JNI_ABORT);
    return rc;
}

/* OceanStage.curlForcing then withBalance on the device; depth null takes the 500 m layer. */
JNIEXPORT jint JNICALL Java_com_cartogenesis_desktop_HipOcean_nPlanSetStress(JNIEnv *env, jclass k, jlong plan,
                                                                             jdouble width_m, jdouble height_m,
                                                                             jdoubleArray eastward,
                                                                             jdoubleArray northward,
                                                                             jdoubleArray depth) {
    (void)k;
    jdouble *e = (*env)->GetDoubleArrayElements(env, eastward, NULL);
    jdouble *n = (*env)->GetDoubleArrayElements(env, northward, NULL);
    jdouble *d = depth ? (*env)->GetDoubleArrayElements(env, depth, NULL) : NULL;
    int rc = cg_ocean_plan_set_stress(PLAN(plan), width_m, height_m, e, n, d);
    if (d) (*env)->ReleaseDoubleArrayElements(env, depth, d, JNI_ABORT);
    (*env)->ReleaseDoubleArrayElements(env, northward, n, JNI_ABORT);
    (*env)->ReleaseDoubleArrayElements(env, eastward, e, JNI_ABORT);
    return rc;
}

JNIEXPORT jint JNICALL Java_com_cartogenesis_desktop_HipOcean_nPlanSetBodies(JNIEnv *env, jclass k, jlong plan,
                                                                             jintArray body_of, jint count) {
    (void)k;
    if (!body_of || count == 0) return cg_ocean_plan_set_bodies(PLAN(plan), NULL, 0);
    jint *b = (*env)->GetIntArrayElements(env, body_of, NULL);
    int rc = cg_ocean_plan_set_bodies(PLAN(plan), (const int32_t *)b, count);
    (*env)->ReleaseIntArrayElements(env, body_of, b, JNI_ABORT);
    return rc;
}

JNIEXPORT jint JNICALL Java_com_cartogenesis_desktop_HipOcean_nPlanSetStart(JNIEnv *env, jclass k, jlong plan,
                                                                            jfloatArray start) {
    (void)k;
    jfloat *s = (*env)->GetFloatArrayElements(env, start, NULL);
    int rc = cg_ocean_plan_set_start(PLAN(plan), s);
    (*env)->ReleaseFloatArrayElements(env, start, s, JNI_ABORT);
    return rc;
}

JNIEXPORT jint JNICALL Java_com_cartogenesis_desktop_HipOcean_nPlanSolve(JNIEnv *env, jclass k, jlong plan,
                                                                         jdouble tol, jint method, jboolean warm,
                                                                         jdoubleArray info) {
    (void)k;
    cg_solution sol = {0, 0.0, 0, 0};
    int rc = cg_ocean_plan_solve(PLAN(plan), tol, method, warm ? 1 : 0, &sol);
    put_info(env, info, &sol);
    return rc;
}

JNIEXPORT jint JNICALL Java_com_cartogenesis_desktop_HipOcean_nPlanRead(JNIEnv *env, jclass k, jlong plan,
                                                                        jfloatArray out) {
    (void)k;
    jfloat *o = (*env)->GetFloatArrayElements(env, out, NULL);
    int rc = cg_ocean_plan_read(PLAN(plan), NULL, o);
    (*env)->ReleaseFloatArrayElements(env, out, o, rc == CG_OK ? 0 : JNI_ABORT);
    return rc;
}

JNIEXPORT jint JNICALL Java_com_cartogenesis_desktop_HipOcean_nPlanVelocities(JNIEnv *env, jclass k, jlong plan,
                                                                              jdouble width_m, jdouble height_m,
                                                                              jfloatArray eastward,
                                                                              jfloatArray northward) {
    (void)k;
    jfloat *e = (*env)->GetFloatArrayElements(env, eastward, NULL);
    jfloat *n = (*env)->GetFloatArrayElements(env, northward, NULL);
    int rc = cg_ocean_plan_velocities(PLAN(plan), width_m, height_m, e, n);
    (*env)->ReleaseFloatArrayElements(env, northward, n, rc == CG_OK ? 0 : JNI_ABORT);
    (*env)->ReleaseFloatArrayElements(env, eastward, e, rc == CG_OK ? 0 : JNI_ABORT);
    return rc;
}

JNIEXPORT jint JNICALL Java_com_cartogenesis_desktop_HipOcean_nPlanSetCoarseNet(JNIEnv *env, jclass k, jlong plan,
                                                                                jint level, jint base,
                                                                                jshortArray weights,
                                                                                jfloatArray depth, jdouble out_scale) {
    (void)k;
    if (!weights) return cg_ocean_plan_set_coarse_net(PLAN(plan), level, base, NULL, 0, NULL, out_scale);
    jsize n = (*env)->GetArrayLength(env, weights);
    jshort *w = (*env)->GetShortArrayElements(env, weights, NULL);
    jfloat *d = depth ? (*env)->GetFloatArrayElements(env, depth, NULL) : NULL;
    int rc = cg_ocean_plan_set_coarse_net(PLAN(plan), level, base, (const uint16_t *)w, (size_t)n, d, out_scale);
    if (d) (*env)->ReleaseFloatArrayElements(env, depth, d, JNI_ABORT);
    (*env)->ReleaseShortArrayElements(env, weights, w, JNI_ABORT);
    return rc;
}
