/*
 * A fake JVM environment for testing cartogenesis_amd/jni/cgocean_jni.c without a JDK: Java arrays,
 * strings, object arrays and direct buffers as plain C objects, and a JNIEnv (tests/jni/jni.h)
 * whose Get<Type>ArrayElements hand out a fresh copy of the array (as a JVM may) and whose
 * Release<Type>ArrayElements copy it back unless the mode is JNI_ABORT, then free it. Outstanding
 * copies are counted, so a test sees an element array the shim never released, an output the shim
 * forgot to copy back and an input it wrote (the write is lost on JNI_ABORT). The fj_* functions
 * are what the tests call through ctypes.
 */
#include <stdlib.h>
#include <string.h>

#include "jni.h"

struct _jobject {
    char kind;    /* 'B','S','I','J','F','D' arrays, 'L' object array, 's' string, 'b' direct buffer */
    jsize len;
    size_t elem;  /* bytes per element */
    void *data;
};

static int outstanding = 0;
static int bad_releases = 0;

static size_t elem_size(char kind) {
    switch (kind) {
        case 'B': return 1;
        case 'S': return 2;
        case 'I': return 4;
        case 'J': return 8;
        case 'F': return 4;
        case 'D': return 8;
        case 'L': return sizeof(jobject);
        default: return 1;
    }
}

__attribute__((visibility("default"))) jobject fj_array(char kind, jsize len) {
    jobject o = calloc(1, sizeof *o);
    o->kind = kind;
    o->len = len;
    o->elem = elem_size(kind);
    o->data = calloc(len > 0 ? (size_t)len : 1, o->elem);
    return o;
}
__attribute__((visibility("default"))) void *fj_data(jobject o) { return o ? o->data : NULL; }
__attribute__((visibility("default"))) jsize fj_len(jobject o) { return o ? o->len : -1; }
__attribute__((visibility("default"))) char fj_kind(jobject o) { return o ? o->kind : 0; }
__attribute__((visibility("default"))) jobject fj_object_at(jobject arr, jsize i) {
    return (arr && arr->kind == 'L' && i >= 0 && i < arr->len) ? ((jobject *)arr->data)[i] : NULL;
}
__attribute__((visibility("default"))) const char *fj_string(jobject s) {
    return (s && s->kind == 's') ? (const char *)s->data : NULL;
}
__attribute__((visibility("default"))) jobject fj_direct_buffer(void *addr, jsize bytes) {
    jobject o = calloc(1, sizeof *o);
    o->kind = 'b';
    o->len = bytes;
    o->elem = 1;
    o->data = addr;
    return o;
}
__attribute__((visibility("default"))) void fj_free(jobject o) {
    if (!o) return;
    if (o->kind != 'b') free(o->data);
    free(o);
}
__attribute__((visibility("default"))) int fj_outstanding(void) { return outstanding; }
__attribute__((visibility("default"))) int fj_bad_releases(void) { return bad_releases; }

static jsize get_length(JNIEnv *env, jarray a) {
    (void)env;
    return a ? a->len : 0;
}
static jstring new_string(JNIEnv *env, const char *utf) {
    (void)env;
    jobject o = calloc(1, sizeof *o);
    o->kind = 's';
    o->len = (jsize)strlen(utf);
    o->elem = 1;
    o->data = malloc((size_t)o->len + 1);
    memcpy(o->data, utf, (size_t)o->len + 1);
    return o;
}
static void set_object(JNIEnv *env, jobjectArray a, jsize i, jobject v) {
    (void)env;
    if (a && a->kind == 'L' && i >= 0 && i < a->len) ((jobject *)a->data)[i] = v;
}
static void *buffer_address(JNIEnv *env, jobject b) {
    (void)env;
    return (b && b->kind == 'b') ? b->data : NULL;
}
static void set_doubles(JNIEnv *env, jdoubleArray a, jsize start, jsize len, const jdouble *buf) {
    (void)env;
    if (a && a->kind == 'D' && start >= 0 && start + len <= a->len) memcpy((jdouble *)a->data + start, buf, sizeof(jdouble) * (size_t)len);
}

/* Element copies: a header before each copy remembers its array, so a release of the wrong array is seen. */
struct copy_head {
    jobject of;
    size_t bytes;
    double pad;
};
static void *get_elems(jarray a, char kind, jboolean *is_copy) {
    if (!a || a->kind != kind) {
        bad_releases++;
        return NULL;
    }
    const size_t bytes = (size_t)a->len * a->elem;
    struct copy_head *h = malloc(sizeof *h + (bytes ? bytes : 1));
    h->of = a;
    h->bytes = bytes;
    memcpy(h + 1, a->data, bytes);
    if (is_copy) *is_copy = JNI_TRUE;
    outstanding++;
    return h + 1;
}
static void release_elems(jarray a, void *elems, jint mode) {
    if (!elems) return;
    struct copy_head *h = (struct copy_head *)elems - 1;
    if (h->of != a) bad_releases++;
    if (mode != JNI_ABORT && h->of) memcpy(h->of->data, elems, h->bytes);
    if (mode != JNI_COMMIT) {
        free(h);
        outstanding--;
    }
}

#define ELEMS(Type, jtype, K)                                                                          \
    static jtype *get_##Type(JNIEnv *env, jarray a, jboolean *c) {                                     \
        (void)env;                                                                                     \
        return (jtype *)get_elems(a, K, c);                                                            \
    }                                                                                                  \
    static void release_##Type(JNIEnv *env, jarray a, jtype *e, jint mode) {                           \
        (void)env;                                                                                     \
        release_elems(a, e, mode);                                                                     \
    }
ELEMS(byte, jbyte, 'B')
ELEMS(short, jshort, 'S')
ELEMS(int, jint, 'I')
ELEMS(long, jlong, 'J')
ELEMS(float, jfloat, 'F')
ELEMS(double, jdouble, 'D')

static const struct JNINativeInterface_ table = {
    get_length, new_string, set_object, buffer_address, set_doubles,
    get_byte, get_short, get_int, get_long, get_float, get_double,
    release_byte, release_short, release_int, release_long, release_float, release_double,
};
static JNIEnv env_ptr = &table;

__attribute__((visibility("default"))) JNIEnv *fj_env(void) { return &env_ptr; }
