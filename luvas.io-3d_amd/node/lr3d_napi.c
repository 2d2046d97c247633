/*
 * lr3d.node — thin N-API binding of the lr3d C ABI (include/lr3d.h) for the TypeScript host.
 *
 * It binds the calls the reference's stage objects make on the GPU — DenseStereo
 * (frontend/src/app/vision/mvs/dense-stereo.ts) and TsdfVolume
 * (frontend/src/app/vision/fusion/tsdf-volume.ts) — and the rest of the C ABI (adjoints, meshing,
 * fusion, launch-shape tuning, counting and timing).  Handles are N-API externals; long calls
 * (solves, filter, integration, tuning, reads) run as async work on the libuv pool and return
 * Promises, the way the reference awaits queue submissions.  Inputs are copied before the call
 * returns.
 */
#define NAPI_VERSION 6
#include <node_api.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/lr3d.h"

#define NAPI_CALL(env, call)                                                                       \
    do {                                                                                           \
        if ((call) != napi_ok) {                                                                   \
            napi_throw_error((env), NULL, "lr3d.node: N-API call failed: " #call);               \
            return NULL;                                                                           \
        }                                                                                          \
    } while (0)

static napi_value throw_status(napi_env env, int32_t status) {
    char msg[600];
    snprintf(msg, sizeof(msg), "lr3d error %d: %s", (int)status, lr3d_last_error());
    napi_throw_error(env, NULL, msg);
    return NULL;
}

static napi_value undefined(napi_env env) {
    napi_value u;
    napi_get_undefined(env, &u);
    return u;
}

/* ---- argument helpers ------------------------------------------------------------------------ */

static void *get_external(napi_env env, napi_value v) {
    void *p = NULL;
    if (napi_get_value_external(env, v, &p) != napi_ok) return NULL;
    return p;
}

static double get_num(napi_env env, napi_value obj, const char *name, double dflt) {
    bool has = false;
    napi_has_named_property(env, obj, name, &has);
    if (!has) return dflt;
    napi_value v;
    double d = dflt;
    napi_get_named_property(env, obj, name, &v);
    napi_valuetype t;
    napi_typeof(env, v, &t);
    if (t == napi_number) napi_get_value_double(env, v, &d);
    else if (t == napi_boolean) {
        bool b;
        napi_get_value_bool(env, v, &b);
        d = b ? 1.0 : 0.0;
    }
    return d;
}

static int get_doubles(napi_env env, napi_value obj, const char *name, double *out, uint32_t n) {
    napi_value arr;
    if (napi_get_named_property(env, obj, name, &arr) != napi_ok) return 0;
    for (uint32_t i = 0; i < n; i++) {
        napi_value e;
        if (napi_get_element(env, arr, i, &e) != napi_ok) return 0;
        if (napi_get_value_double(env, e, &out[i]) != napi_ok) return 0;
    }
    return 1;
}

/* Copy of a typed array of element type `want` (NULL when the value is null/undefined).  *bad is
 * set when the value is anything else than such an array (another element type included), so the
 * caller can throw instead of reading the wrong number of bytes. */
static void *copy_typed(napi_env env, napi_value v, napi_typedarray_type want, size_t *count, int *bad) {
    *bad = 0;
    if (count) *count = 0;
    napi_valuetype t;
    napi_typeof(env, v, &t);
    if (t == napi_null || t == napi_undefined) return NULL;
    bool is_ta = false;
    napi_is_typedarray(env, v, &is_ta);
    napi_typedarray_type type;
    size_t length = 0, offset = 0;
    void *data = NULL;
    napi_value ab;
    if (!is_ta || napi_get_typedarray_info(env, v, &type, &length, &data, &ab, &offset) != napi_ok ||
        type != want) {
        *bad = 1;
        return NULL;
    }
    size_t elem = (want == napi_uint8_array) ? 1 : (want == napi_float64_array) ? 8 : 4;
    void *copy = malloc(length * elem + 1);
    if (!copy) {
        *bad = 1;
        return NULL;
    }
    if (length) memcpy(copy, data, length * elem);
    if (count) *count = length;
    return copy;
}

static int read_view(napi_env env, napi_value o, lr3d_view *v) {
    memset(v, 0, sizeof(*v));
    v->width = (int32_t)get_num(env, o, "width", 0);
    v->height = (int32_t)get_num(env, o, "height", 0);
    v->fx = get_num(env, o, "fx", 0);
    v->fy = get_num(env, o, "fy", 0);
    v->cx = get_num(env, o, "cx", 0);
    v->cy = get_num(env, o, "cy", 0);
    return get_doubles(env, o, "R", v->R, 9) && get_doubles(env, o, "t", v->t, 3);
}

/* ---- async work -------------------------------------------------------------------------------- */

enum { OP_SOLVE, OP_SOLVE_BATCHED, OP_FILTER, OP_INTEGRATE, OP_READ_DEPTH, OP_READ_FILTERED,
       OP_READ_HYP, OP_READ_TSDF, OP_NCC_BACKWARD, OP_NCC_REFINE, OP_TSDF_BACKWARD, OP_MESH, OP_FUSE,
       OP_SOLVE_NOALLOC, OP_STEREO_AUTOTUNE, OP_TSDF_AUTOTUNE, OP_INTEGRATE_HOST };

typedef struct {
    int op;
    napi_deferred deferred;
    napi_async_work work;
    int32_t status;
    char error[600];
    lr3d_stereo *st;
    lr3d_tsdf *ts;
    lr3d_solve_desc *descs;
    float **priors;
    int32_t n;
    int32_t layer, map, needed, n_checks;
    int32_t checks[LR3D_MAX_SOURCES];
    int32_t *layers;
    lr3d_integration_view *iviews;
    void *out_a, *out_b;
    size_t bytes_a, bytes_b;
    /* adjoints */
    float *weights, *targets;
    uint32_t *voxels;
    int64_t n_voxels;
    float step_depth, step_normal;
    void *out_c, *out_d, *out_e;
    size_t bytes_c, bytes_d, bytes_e;
    double loss;
    /* mesh extraction */
    double min_weight, fraction;
    int32_t min_faces;
    /* dense-cloud fusion */
    lr3d_thumb *thumbs;
    uint8_t **rgbs;
    double cell, final_cell;
    int64_t cell_budget;
    int32_t min_support;
    /* tuning, host-map integration */
    int32_t repeats;
    uint32_t *packed;
} job_t;

static void job_free(job_t *j) {
    if (j->priors) {
        for (int32_t i = 0; i < j->n; i++) free(j->priors[i]);
        free(j->priors);
    }
    free(j->descs);
    free(j->layers);
    free(j->iviews);
    free(j->out_a);
    free(j->out_b);
    free(j->out_c);
    free(j->out_d);
    free(j->out_e);
    free(j->weights);
    free(j->targets);
    free(j->voxels);
    free(j->packed);
    if (j->rgbs) {
        for (int32_t i = 0; i < j->n; i++) free(j->rgbs[i]);
        free(j->rgbs);
    }
    free(j->thumbs);
    free(j);
}

static void job_execute(napi_env env, void *data) {
    (void)env;
    job_t *j = (job_t *)data;
    switch (j->op) {
    case OP_SOLVE: j->status = lr3d_stereo_solve(j->st, &j->descs[0]); break;
    case OP_SOLVE_BATCHED: j->status = lr3d_stereo_solve_batched(j->st, j->descs, j->n); break;
    case OP_SOLVE_NOALLOC: j->status = lr3d_stereo_solve_batched_noalloc(j->st, j->descs, j->n); break;
    case OP_STEREO_AUTOTUNE: j->status = lr3d_stereo_autotune(j->st, j->descs, j->n, j->repeats); break;
    case OP_TSDF_AUTOTUNE:
        j->status = lr3d_tsdf_autotune(j->ts, j->st, j->layers, j->map, j->iviews, j->n, j->repeats);
        break;
    case OP_INTEGRATE_HOST: j->status = lr3d_tsdf_integrate_host(j->ts, j->packed, &j->iviews[0]); break;
    case OP_FILTER: j->status = lr3d_stereo_filter(j->st, j->layer, j->checks, j->n_checks, j->needed); break;
    case OP_INTEGRATE:
        j->status = lr3d_tsdf_integrate_batched(j->ts, j->st, j->layers, j->map, j->iviews, j->n);
        break;
    case OP_READ_DEPTH: j->status = lr3d_stereo_read_depth(j->st, j->layer, (uint32_t *)j->out_a); break;
    case OP_READ_FILTERED: j->status = lr3d_stereo_read_filtered(j->st, j->layer, (uint32_t *)j->out_a); break;
    case OP_READ_HYP: j->status = lr3d_stereo_read_hyp(j->st, j->layer, (float *)j->out_a, (float *)j->out_b); break;
    case OP_READ_TSDF: j->status = lr3d_tsdf_read(j->ts, (float *)j->out_a, (float *)j->out_b); break;
    case OP_NCC_BACKWARD:
        j->status = lr3d_ncc_backward(j->st, &j->descs[0], j->weights, (float *)j->out_a, (float *)j->out_b,
                                      (float *)j->out_c, (double *)j->out_d, &j->loss);
        break;
    case OP_NCC_REFINE:
        j->status = lr3d_ncc_refine_step(j->st, j->descs, j->n, j->step_depth, j->step_normal, (double *)j->out_a,
                                         (double *)j->out_b, (double *)j->out_c);
        break;
    case OP_MESH: {
        /* extractSurface, then removeSmallComponents when fraction >= 0 (tsdf-meshing.ts:84-90) */
        lr3d_mesh *m = NULL, *c = NULL;
        j->status = lr3d_tsdf_extract_surface(j->ts, j->min_weight, &m);
        if (j->status == LR3D_OK && j->fraction >= 0.0) {
            j->status = lr3d_mesh_remove_small_components(m, j->fraction, j->min_faces, &c);
            lr3d_mesh_destroy(m);
            m = c;
        }
        int64_t nv = 0, ni = 0;
        if (j->status == LR3D_OK) j->status = lr3d_mesh_size(m, &nv, &ni);
        if (j->status == LR3D_OK) {
            j->bytes_a = j->bytes_b = (size_t)nv * 12u;
            j->bytes_c = (size_t)ni * 4u;
            j->out_a = malloc(j->bytes_a + 1);
            j->out_b = malloc(j->bytes_b + 1);
            j->out_c = malloc(j->bytes_c + 1);
            j->status = lr3d_mesh_read(m, (float *)j->out_a, (float *)j->out_b, (uint32_t *)j->out_c);
        }
        lr3d_mesh_destroy(m);
        break;
    }
    case OP_FUSE: {
        /* fuseDepthMaps (depth-fusion.ts:23-74): positions, normals, colors, support */
        lr3d_cloud *c = NULL;
        j->status = lr3d_fuse_depth_maps(j->st, j->layers, j->map, j->n, j->thumbs, j->cell, j->cell_budget,
                                         j->min_support, &c, &j->final_cell);
        int64_t np = 0;
        if (j->status == LR3D_OK) j->status = lr3d_cloud_size(c, &np);
        if (j->status == LR3D_OK) {
            j->bytes_a = j->bytes_b = (size_t)np * 12u;
            j->bytes_c = (size_t)np * 3u;
            j->bytes_d = (size_t)np * 2u;
            j->out_a = malloc(j->bytes_a + 1);
            j->out_b = malloc(j->bytes_b + 1);
            j->out_c = malloc(j->bytes_c + 1);
            j->out_d = malloc(j->bytes_d + 1);
            j->status = lr3d_cloud_read(c, (float *)j->out_a, (float *)j->out_b, (uint8_t *)j->out_c,
                                        (uint16_t *)j->out_d);
        }
        lr3d_cloud_destroy(c);
        break;
    }
    case OP_TSDF_BACKWARD:
        j->status = lr3d_tsdf_backward(j->ts, j->st, j->layers, j->map, j->iviews, j->n, j->voxels, j->targets,
                                       j->n_voxels, (float *)j->out_a, (float *)j->out_b, (double *)j->out_c, &j->loss);
        break;
    }
    if (j->status != LR3D_OK) snprintf(j->error, sizeof(j->error), "lr3d error %d: %s", (int)j->status, lr3d_last_error());
}

static napi_value typed_from(napi_env env, void *src, size_t bytes, napi_typedarray_type type, size_t elem) {
    napi_value ab, ta;
    void *dst = NULL;
    napi_create_arraybuffer(env, bytes, &dst, &ab);
    if (bytes) memcpy(dst, src, bytes);
    napi_create_typedarray(env, type, bytes / elem, ab, 0, &ta);
    return ta;
}

static void job_complete(napi_env env, napi_status status, void *data) {
    job_t *j = (job_t *)data;
    if (status != napi_ok || j->status != LR3D_OK) {
        napi_value msg, err;
        napi_create_string_utf8(env, j->status != LR3D_OK ? j->error : "lr3d.node: async work cancelled",
                                NAPI_AUTO_LENGTH, &msg);
        napi_create_error(env, NULL, msg, &err);
        napi_reject_deferred(env, j->deferred, err);
    } else {
        napi_value result = undefined(env);
        if (j->op == OP_READ_DEPTH || j->op == OP_READ_FILTERED) {
            result = typed_from(env, j->out_a, j->bytes_a, napi_uint32_array, 4);
        } else if (j->op == OP_NCC_BACKWARD) {
            napi_value loss;
            napi_create_object(env, &result);
            napi_set_named_property(env, result, "cost", typed_from(env, j->out_a, j->bytes_a, napi_float32_array, 4));
            napi_set_named_property(env, result, "gradDepth", typed_from(env, j->out_b, j->bytes_b, napi_float32_array, 4));
            napi_set_named_property(env, result, "gradNormal", typed_from(env, j->out_c, j->bytes_c, napi_float32_array, 4));
            napi_set_named_property(env, result, "gradPose", typed_from(env, j->out_d, j->bytes_d, napi_float64_array, 8));
            napi_create_double(env, j->loss, &loss);
            napi_set_named_property(env, result, "loss", loss);
        } else if (j->op == OP_NCC_REFINE) {
            napi_create_object(env, &result);
            napi_set_named_property(env, result, "gradPose", typed_from(env, j->out_a, j->bytes_a, napi_float64_array, 8));
            napi_set_named_property(env, result, "lossBefore", typed_from(env, j->out_b, j->bytes_b, napi_float64_array, 8));
            napi_set_named_property(env, result, "lossAfter", typed_from(env, j->out_c, j->bytes_c, napi_float64_array, 8));
        } else if (j->op == OP_MESH) {
            napi_create_object(env, &result);
            napi_set_named_property(env, result, "positions", typed_from(env, j->out_a, j->bytes_a, napi_float32_array, 4));
            napi_set_named_property(env, result, "normals", typed_from(env, j->out_b, j->bytes_b, napi_float32_array, 4));
            napi_set_named_property(env, result, "indices", typed_from(env, j->out_c, j->bytes_c, napi_uint32_array, 4));
        } else if (j->op == OP_FUSE) {
            napi_value cell;
            napi_create_object(env, &result);
            napi_set_named_property(env, result, "positions", typed_from(env, j->out_a, j->bytes_a, napi_float32_array, 4));
            napi_set_named_property(env, result, "normals", typed_from(env, j->out_b, j->bytes_b, napi_float32_array, 4));
            napi_set_named_property(env, result, "colors", typed_from(env, j->out_c, j->bytes_c, napi_uint8_array, 1));
            napi_set_named_property(env, result, "support", typed_from(env, j->out_d, j->bytes_d, napi_uint16_array, 2));
            napi_create_double(env, j->final_cell, &cell);
            napi_set_named_property(env, result, "cell", cell);
        } else if (j->op == OP_TSDF_BACKWARD) {
            napi_value loss;
            napi_create_object(env, &result);
            napi_set_named_property(env, result, "tsdf", typed_from(env, j->out_a, j->bytes_a, napi_float32_array, 4));
            napi_set_named_property(env, result, "gradDepth", typed_from(env, j->out_b, j->bytes_b, napi_float32_array, 4));
            napi_set_named_property(env, result, "gradPose", typed_from(env, j->out_c, j->bytes_c, napi_float64_array, 8));
            napi_create_double(env, j->loss, &loss);
            napi_set_named_property(env, result, "loss", loss);
        } else if (j->op == OP_READ_HYP || j->op == OP_READ_TSDF) {
            napi_create_object(env, &result);
            napi_set_named_property(env, result, j->op == OP_READ_HYP ? "hyp" : "tsdf",
                                    typed_from(env, j->out_a, j->bytes_a, napi_float32_array, 4));
            napi_set_named_property(env, result, j->op == OP_READ_HYP ? "cost" : "weight",
                                    typed_from(env, j->out_b, j->bytes_b, napi_float32_array, 4));
        }
        napi_resolve_deferred(env, j->deferred, result);
    }
    napi_delete_async_work(env, j->work);
    job_free(j);
}

static napi_value queue(napi_env env, job_t *j) {
    napi_value promise, name;
    NAPI_CALL(env, napi_create_promise(env, &j->deferred, &promise));
    NAPI_CALL(env, napi_create_string_utf8(env, "lr3d", NAPI_AUTO_LENGTH, &name));
    NAPI_CALL(env, napi_create_async_work(env, NULL, name, job_execute, job_complete, j, &j->work));
    NAPI_CALL(env, napi_queue_async_work(env, j->work));
    return promise;
}

static job_t *new_job(int op) {
    job_t *j = (job_t *)calloc(1, sizeof(job_t));
    if (j) j->op = op;
    return j;
}

/* ---- bindings ------------------------------------------------------------------------------------ */

#define ARGS(n)                                                                                    \
    size_t argc = (n);                                                                             \
    napi_value argv[(n) > 0 ? (n) : 1];                                                            \
    NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL))

static napi_value ctx_create(napi_env env, napi_callback_info info) {
    ARGS(1);
    int32_t device = 0;
    if (argc > 0) napi_get_value_int32(env, argv[0], &device);
    lr3d_ctx *ctx = NULL;
    int32_t rc = lr3d_ctx_create(device, &ctx);
    if (rc != LR3D_OK) return throw_status(env, rc);
    napi_value out;
    NAPI_CALL(env, napi_create_external(env, ctx, NULL, NULL, &out));
    return out;
}

static napi_value ctx_destroy(napi_env env, napi_callback_info info) {
    ARGS(1);
    lr3d_ctx_destroy((lr3d_ctx *)get_external(env, argv[0]));
    return undefined(env);
}

/* stereoCreate(ctx, views[], flags?) (flags: LR3D_STEREO_PRIOR_POOL = 1) */
static napi_value stereo_create(napi_env env, napi_callback_info info) {
    ARGS(3);
    lr3d_ctx *ctx = (lr3d_ctx *)get_external(env, argv[0]);
    uint32_t flags = 0;
    if (argc > 2) napi_get_value_uint32(env, argv[2], &flags);
    uint32_t n = 0;
    NAPI_CALL(env, napi_get_array_length(env, argv[1], &n));
    lr3d_view *views = (lr3d_view *)calloc(n ? n : 1, sizeof(lr3d_view));
    for (uint32_t i = 0; i < n; i++) {
        napi_value o;
        napi_get_element(env, argv[1], i, &o);
        if (!read_view(env, o, &views[i])) {
            free(views);
            napi_throw_type_error(env, NULL, "lr3d.node: a view needs width, height, fx, fy, cx, cy, R[9], t[3]");
            return NULL;
        }
    }
    lr3d_stereo *st = NULL;
    int32_t rc = lr3d_stereo_create(ctx, views, (int32_t)n, flags, &st);
    free(views);
    if (rc != LR3D_OK) return throw_status(env, rc);
    napi_value out;
    NAPI_CALL(env, napi_create_external(env, st, NULL, NULL, &out));
    return out;
}

static napi_value stereo_destroy(napi_env env, napi_callback_info info) {
    ARGS(1);
    lr3d_stereo_destroy((lr3d_stereo *)get_external(env, argv[0]));
    return undefined(env);
}

static napi_value stereo_set_image(napi_env env, napi_callback_info info) {
    ARGS(5);
    lr3d_stereo *st = (lr3d_stereo *)get_external(env, argv[0]);
    int32_t layer, w, h;
    napi_get_value_int32(env, argv[1], &layer);
    napi_get_value_int32(env, argv[3], &w);
    napi_get_value_int32(env, argv[4], &h);
    size_t count = 0;
    int bad = 0;
    uint8_t *gray = (uint8_t *)copy_typed(env, argv[2], napi_uint8_array, &count, &bad);
    if (!gray || count != (size_t)w * (size_t)h) {
        free(gray);
        napi_throw_type_error(env, NULL, "lr3d.node: setImage needs a Uint8Array of width*height bytes");
        return NULL;
    }
    int32_t rc = lr3d_stereo_set_image(st, layer, gray, w, h);
    free(gray);
    if (rc != LR3D_OK) return throw_status(env, rc);
    return undefined(env);
}

/* SolveOptions (DS:16-29) + ref -> lr3d_solve_desc; the prior is copied into *prior_out.  Returns 1,
 * 0 when ref/sources are missing, -1 when the prior is not a Float32Array of `pixels` (the
 * reference layer's width*height, set by lr3d.js) elements. */
static int read_solve(napi_env env, napi_value o, lr3d_solve_desc *d, float **prior_out) {
    memset(d, 0, sizeof(*d));
    d->ref = (int32_t)get_num(env, o, "ref", 0);
    napi_value sources;
    if (napi_get_named_property(env, o, "sources", &sources) != napi_ok) return 0;
    uint32_t n = 0;
    napi_get_array_length(env, sources, &n);
    d->n_sources = n > LR3D_MAX_SOURCES ? LR3D_MAX_SOURCES : (int32_t)n;
    for (int32_t i = 0; i < d->n_sources; i++) {
        napi_value e;
        napi_get_element(env, sources, (uint32_t)i, &e);
        napi_get_value_int32(env, e, &d->sources[i]);
    }
    d->min_depth = get_num(env, o, "minDepth", 0);
    d->max_depth = get_num(env, o, "maxDepth", 0);
    d->geometric = (int32_t)get_num(env, o, "geometric", 0);
    d->iterations = (int32_t)get_num(env, o, "iterations", 5);
    d->from_coarse = (int32_t)get_num(env, o, "fromCoarse", 0);
    d->pack = (int32_t)get_num(env, o, "pack", 1);
    d->topk = (int32_t)get_num(env, o, "topk", 0);
    *prior_out = NULL;
    bool has = false;
    napi_has_named_property(env, o, "prior", &has);
    if (has) {
        napi_value p;
        napi_get_named_property(env, o, "prior", &p);
        size_t count = 0;
        int bad = 0;
        *prior_out = (float *)copy_typed(env, p, napi_float32_array, &count, &bad);
        if (bad) return -1;
        if (*prior_out && (double)count != get_num(env, o, "pixels", -1)) return -1;
    }
    d->prior = *prior_out;
    return 1;
}

/* Solve descriptors of argv[1] (one object or an array) into a new job of `op` (OP_SOLVE for a
 * single object given to a batched op).  NULL after throwing. */
static job_t *solve_job(napi_env env, napi_value *argv, int op) {
    job_t *j = new_job(op);
    j->st = (lr3d_stereo *)get_external(env, argv[0]);
    bool is_array = false;
    napi_is_array(env, argv[1], &is_array);
    uint32_t n = 1;
    if (is_array) napi_get_array_length(env, argv[1], &n);
    else if (op == OP_SOLVE_BATCHED) j->op = OP_SOLVE;
    j->n = (int32_t)n;
    j->descs = (lr3d_solve_desc *)calloc(n ? n : 1, sizeof(lr3d_solve_desc));
    j->priors = (float **)calloc(n ? n : 1, sizeof(float *));
    for (uint32_t i = 0; i < n; i++) {
        napi_value o = argv[1];
        if (is_array) napi_get_element(env, argv[1], i, &o);
        const int ok = read_solve(env, o, &j->descs[i], &j->priors[i]);
        if (ok <= 0) {
            job_free(j);
            if (ok < 0) napi_throw_range_error(env, NULL, "lr3d.node: prior must be a Float32Array of width*height");
            else napi_throw_type_error(env, NULL, "lr3d.node: solve options need ref and sources");
            return NULL;
        }
    }
    return j;
}

static napi_value stereo_solve(napi_env env, napi_callback_info info) {
    ARGS(2);
    job_t *j = solve_job(env, argv, OP_SOLVE_BATCHED);
    return j ? queue(env, j) : NULL;
}

/* stereoSolveNoalloc(stereo, solves[]): lr3d_stereo_solve_batched_noalloc */
static napi_value stereo_solve_noalloc(napi_env env, napi_callback_info info) {
    ARGS(2);
    job_t *j = solve_job(env, argv, OP_SOLVE_NOALLOC);
    return j ? queue(env, j) : NULL;
}

/* stereoAutotune(stereo, solves[], repeats) */
static napi_value stereo_autotune(napi_env env, napi_callback_info info) {
    ARGS(3);
    job_t *j = solve_job(env, argv, OP_STEREO_AUTOTUNE);
    if (!j) return NULL;
    j->repeats = 3;
    if (argc > 2) napi_get_value_int32(env, argv[2], &j->repeats);
    return queue(env, j);
}

static napi_value stereo_filter(napi_env env, napi_callback_info info) {
    ARGS(4);
    job_t *j = new_job(OP_FILTER);
    j->st = (lr3d_stereo *)get_external(env, argv[0]);
    napi_get_value_int32(env, argv[1], &j->layer);
    uint32_t n = 0;
    napi_get_array_length(env, argv[2], &n);
    j->n_checks = n > LR3D_MAX_SOURCES ? LR3D_MAX_SOURCES : (int32_t)n;
    for (int32_t i = 0; i < j->n_checks; i++) {
        napi_value e;
        napi_get_element(env, argv[2], (uint32_t)i, &e);
        napi_get_value_int32(env, e, &j->checks[i]);
    }
    napi_get_value_int32(env, argv[3], &j->needed);
    return queue(env, j);
}

static napi_value stereo_read(napi_env env, napi_callback_info info, int op) {
    ARGS(4);
    job_t *j = new_job(op);
    j->st = (lr3d_stereo *)get_external(env, argv[0]);
    napi_get_value_int32(env, argv[1], &j->layer);
    int32_t w = 0, h = 0;
    napi_get_value_int32(env, argv[2], &w);
    napi_get_value_int32(env, argv[3], &h);
    size_t px = (size_t)w * (size_t)h;
    j->bytes_a = px * (op == OP_READ_HYP ? 16 : 4);
    j->out_a = malloc(j->bytes_a + 1);
    if (op == OP_READ_HYP) {
        j->bytes_b = px * 4;
        j->out_b = malloc(j->bytes_b + 1);
    }
    return queue(env, j);
}

static napi_value stereo_read_depth(napi_env env, napi_callback_info info) { return stereo_read(env, info, OP_READ_DEPTH); }
static napi_value stereo_read_filtered(napi_env env, napi_callback_info info) { return stereo_read(env, info, OP_READ_FILTERED); }
static napi_value stereo_read_hyp(napi_env env, napi_callback_info info) { return stereo_read(env, info, OP_READ_HYP); }

static napi_value stereo_set_coarse(napi_env env, napi_callback_info info) {
    ARGS(2);
    napi_valuetype t;
    napi_typeof(env, argv[1], &t);
    int32_t rc = lr3d_stereo_set_coarse((lr3d_stereo *)get_external(env, argv[0]),
                                        t == napi_external ? (lr3d_stereo *)get_external(env, argv[1]) : NULL);
    if (rc != LR3D_OK) return throw_status(env, rc);
    return undefined(env);
}

static napi_value tsdf_create(napi_env env, napi_callback_info info) {
    ARGS(6);
    lr3d_ctx *ctx = (lr3d_ctx *)get_external(env, argv[0]);
    double origin[3], voxel = 0, trunc = 4;
    int32_t dims[3];
    for (uint32_t a = 0; a < 3; a++) {
        napi_value e;
        napi_get_element(env, argv[1], a, &e);
        napi_get_value_double(env, e, &origin[a]);
        napi_get_element(env, argv[3], a, &e);
        napi_get_value_int32(env, e, &dims[a]);
    }
    napi_get_value_double(env, argv[2], &voxel);
    size_t count = 0;
    int bad = 0;
    int32_t *coords = (int32_t *)copy_typed(env, argv[4], napi_int32_array, &count, &bad);
    napi_get_value_double(env, argv[5], &trunc);
    if (!coords) {
        napi_throw_type_error(env, NULL, "lr3d.node: coords must be an Int32Array");
        return NULL;
    }
    lr3d_tsdf *ts = NULL;
    int32_t rc = lr3d_tsdf_create(ctx, origin, voxel, dims, coords, (int32_t)(count / 4), trunc, &ts);
    free(coords);
    if (rc != LR3D_OK) return throw_status(env, rc);
    napi_value out;
    NAPI_CALL(env, napi_create_external(env, ts, NULL, NULL, &out));
    return out;
}

static napi_value tsdf_destroy(napi_env env, napi_callback_info info) {
    ARGS(1);
    lr3d_tsdf_destroy((lr3d_tsdf *)get_external(env, argv[0]));
    return undefined(env);
}

/* IntegrationView (TV:67-79) + size -> lr3d_integration_view (prior copied).  Returns 1, 0 when
 * the pose is missing, -1 when the prior is not a Float32Array of width*height elements. */
static int read_iview(napi_env env, napi_value o, lr3d_integration_view *v, float **prior) {
    memset(v, 0, sizeof(*v));
    v->width = (int32_t)get_num(env, o, "width", 0);
    v->height = (int32_t)get_num(env, o, "height", 0);
    v->fx = get_num(env, o, "fx", 0);
    v->fy = get_num(env, o, "fy", 0);
    v->cx = get_num(env, o, "cx", 0);
    v->cy = get_num(env, o, "cy", 0);
    v->max_depth = get_num(env, o, "maxDepth", 0);
    v->confidence = get_num(env, o, "confidence", 0);
    *prior = NULL;
    bool has = false;
    napi_has_named_property(env, o, "prior", &has);
    if (has) {
        napi_value p;
        napi_get_named_property(env, o, "prior", &p);
        size_t count = 0;
        int bad = 0;
        *prior = (float *)copy_typed(env, p, napi_float32_array, &count, &bad);
        if (bad || (*prior && count != (size_t)v->width * (size_t)v->height)) return -1;
    }
    v->prior = *prior;
    return get_doubles(env, o, "R", v->R, 9) && get_doubles(env, o, "t", v->t, 3);
}

/* (volume, stereo, layers[], map, views[]) of argv into a new job of `op`.  NULL after throwing. */
static job_t *integrate_job(napi_env env, napi_value *argv, int op) {
    job_t *j = new_job(op);
    j->ts = (lr3d_tsdf *)get_external(env, argv[0]);
    j->st = (lr3d_stereo *)get_external(env, argv[1]);
    uint32_t n = 0;
    napi_get_array_length(env, argv[2], &n);
    napi_get_value_int32(env, argv[3], &j->map);
    j->n = (int32_t)n;
    j->layers = (int32_t *)calloc(n ? n : 1, sizeof(int32_t));
    j->iviews = (lr3d_integration_view *)calloc(n ? n : 1, sizeof(lr3d_integration_view));
    j->priors = (float **)calloc(n ? n : 1, sizeof(float *));
    for (uint32_t i = 0; i < n; i++) {
        napi_value e, o;
        napi_get_element(env, argv[2], i, &e);
        napi_get_value_int32(env, e, &j->layers[i]);
        napi_get_element(env, argv[4], i, &o);
        const int ok = read_iview(env, o, &j->iviews[i], &j->priors[i]);
        if (ok <= 0) {
            job_free(j);
            if (ok < 0) napi_throw_range_error(env, NULL, "lr3d.node: prior must be a Float32Array of width*height");
            else napi_throw_type_error(env, NULL, "lr3d.node: an integration view needs intrinsics, R[9], t[3]");
            return NULL;
        }
    }
    return j;
}

static napi_value tsdf_integrate(napi_env env, napi_callback_info info) {
    ARGS(5);
    job_t *j = integrate_job(env, argv, OP_INTEGRATE);
    return j ? queue(env, j) : NULL;
}

/* tsdfAutotune(volume, stereo, layers[], map, views[], repeats) */
static napi_value tsdf_autotune(napi_env env, napi_callback_info info) {
    ARGS(6);
    job_t *j = integrate_job(env, argv, OP_TSDF_AUTOTUNE);
    if (!j) return NULL;
    j->repeats = 3;
    if (argc > 5) napi_get_value_int32(env, argv[5], &j->repeats);
    return queue(env, j);
}

/* tsdfIntegrateHost(volume, packed: Uint32Array (width*height), view) */
static napi_value tsdf_integrate_host(napi_env env, napi_callback_info info) {
    ARGS(3);
    job_t *j = new_job(OP_INTEGRATE_HOST);
    j->ts = (lr3d_tsdf *)get_external(env, argv[0]);
    j->n = 1;
    j->iviews = (lr3d_integration_view *)calloc(1, sizeof(lr3d_integration_view));
    j->priors = (float **)calloc(1, sizeof(float *));
    const int ok = read_iview(env, argv[2], &j->iviews[0], &j->priors[0]);
    size_t count = 0;
    int bad = 0;
    j->packed = (uint32_t *)copy_typed(env, argv[1], napi_uint32_array, &count, &bad);
    if (ok <= 0 || bad || !j->packed || count != (size_t)j->iviews[0].width * (size_t)j->iviews[0].height) {
        job_free(j);
        napi_throw_range_error(env, NULL, "lr3d.node: tsdfIntegrateHost needs a Uint32Array of width*height and a view");
        return NULL;
    }
    return queue(env, j);
}

/* ---- synchronous calls (no device work worth a worker thread, or plain uploads) ---------------- */

static napi_value status_or_undefined(napi_env env, int32_t rc) {
    return rc == LR3D_OK ? undefined(env) : throw_status(env, rc);
}

static napi_value ctx_synchronize(napi_env env, napi_callback_info info) {
    ARGS(1);
    return status_or_undefined(env, lr3d_ctx_synchronize((lr3d_ctx *)get_external(env, argv[0])));
}

static napi_value ctx_set_profiling(napi_env env, napi_callback_info info) {
    ARGS(2);
    bool on = false;
    napi_get_value_bool(env, argv[1], &on);
    return status_or_undefined(env, lr3d_ctx_set_profiling((lr3d_ctx *)get_external(env, argv[0]), on ? 1 : 0));
}

/* ctxKernelStats(ctx, name) -> { launches, ms } */
static napi_value ctx_kernel_stats(napi_env env, napi_callback_info info) {
    ARGS(2);
    char name[128];
    size_t len = 0;
    NAPI_CALL(env, napi_get_value_string_utf8(env, argv[1], name, sizeof(name), &len));
    int64_t launches = 0;
    double ms = 0;
    const int32_t rc = lr3d_ctx_kernel_stats((lr3d_ctx *)get_external(env, argv[0]), name, &launches, &ms);
    if (rc != LR3D_OK) return throw_status(env, rc);
    napi_value out, a, b;
    napi_create_object(env, &out);
    napi_create_int64(env, launches, &a);
    napi_create_double(env, ms, &b);
    napi_set_named_property(env, out, "launches", a);
    napi_set_named_property(env, out, "ms", b);
    return out;
}

/* stereoGetTile(stereo, kernel) -> [tx, ty]; stereoSetTile(stereo, kernel, tx, ty) */
static napi_value stereo_get_tile(napi_env env, napi_callback_info info) {
    ARGS(2);
    char kernel[32];
    size_t len = 0;
    NAPI_CALL(env, napi_get_value_string_utf8(env, argv[1], kernel, sizeof(kernel), &len));
    int32_t tx = 0, ty = 0;
    const int32_t rc = lr3d_stereo_get_tile((lr3d_stereo *)get_external(env, argv[0]), kernel, &tx, &ty);
    if (rc != LR3D_OK) return throw_status(env, rc);
    napi_value out, a, b;
    napi_create_array_with_length(env, 2, &out);
    napi_create_int32(env, tx, &a);
    napi_create_int32(env, ty, &b);
    napi_set_element(env, out, 0, a);
    napi_set_element(env, out, 1, b);
    return out;
}

static napi_value stereo_set_tile(napi_env env, napi_callback_info info) {
    ARGS(4);
    char kernel[32];
    size_t len = 0;
    NAPI_CALL(env, napi_get_value_string_utf8(env, argv[1], kernel, sizeof(kernel), &len));
    int32_t tx = 0, ty = 0;
    napi_get_value_int32(env, argv[2], &tx);
    napi_get_value_int32(env, argv[3], &ty);
    return status_or_undefined(env, lr3d_stereo_set_tile((lr3d_stereo *)get_external(env, argv[0]), kernel, tx, ty));
}

/* stereoWriteDepth(stereo, layer, packed: Uint32Array, pixels); stereoWriteHyp(stereo, layer,
 * planes: Float32Array, pixels) (4 floats per pixel) */
static napi_value stereo_write(napi_env env, napi_callback_info info, int hyp) {
    ARGS(4);
    int32_t layer = 0;
    double pixels = 0;
    napi_get_value_int32(env, argv[1], &layer);
    napi_get_value_double(env, argv[3], &pixels);
    size_t count = 0;
    int bad = 0;
    void *data = copy_typed(env, argv[2], hyp ? napi_float32_array : napi_uint32_array, &count, &bad);
    if (bad || !data || (double)count != pixels * (hyp ? 4.0 : 1.0)) {
        free(data);
        napi_throw_range_error(env, NULL, hyp ? "lr3d.node: planes must be a Float32Array of 4*width*height"
                                              : "lr3d.node: packed must be a Uint32Array of width*height");
        return NULL;
    }
    lr3d_stereo *st = (lr3d_stereo *)get_external(env, argv[0]);
    const int32_t rc = hyp ? lr3d_stereo_write_hyp(st, layer, (const float *)data)
                           : lr3d_stereo_write_depth(st, layer, (const uint32_t *)data);
    free(data);
    return status_or_undefined(env, rc);
}
static napi_value stereo_write_depth(napi_env env, napi_callback_info info) { return stereo_write(env, info, 0); }
static napi_value stereo_write_hyp(napi_env env, napi_callback_info info) { return stereo_write(env, info, 1); }

/* tsdfGetBricksPerBlock(volume) -> n; tsdfSetBricksPerBlock(volume, n) */
static napi_value tsdf_get_bricks_per_block(napi_env env, napi_callback_info info) {
    ARGS(1);
    int32_t n = 0;
    const int32_t rc = lr3d_tsdf_get_bricks_per_block((lr3d_tsdf *)get_external(env, argv[0]), &n);
    if (rc != LR3D_OK) return throw_status(env, rc);
    napi_value out;
    napi_create_int32(env, n, &out);
    return out;
}

static napi_value tsdf_set_bricks_per_block(napi_env env, napi_callback_info info) {
    ARGS(2);
    int32_t n = 0;
    napi_get_value_int32(env, argv[1], &n);
    return status_or_undefined(env, lr3d_tsdf_set_bricks_per_block((lr3d_tsdf *)get_external(env, argv[0]), n));
}

/* tsdfGetXcdChunk(volume) -> n; tsdfSetXcdChunk(volume, n) */
static napi_value tsdf_get_xcd_chunk(napi_env env, napi_callback_info info) {
    ARGS(1);
    int32_t n = 0;
    const int32_t rc = lr3d_tsdf_get_xcd_chunk((lr3d_tsdf *)get_external(env, argv[0]), &n);
    if (rc != LR3D_OK) return throw_status(env, rc);
    napi_value out;
    napi_create_int32(env, n, &out);
    return out;
}

static napi_value tsdf_set_xcd_chunk(napi_env env, napi_callback_info info) {
    ARGS(2);
    int32_t n = 0;
    napi_get_value_int32(env, argv[1], &n);
    return status_or_undefined(env, lr3d_tsdf_set_xcd_chunk((lr3d_tsdf *)get_external(env, argv[0]), n));
}

static napi_value tsdf_set_flag(napi_env env, napi_callback_info info, int which) {
    ARGS(2);
    bool on = false;
    napi_get_value_bool(env, argv[1], &on);
    lr3d_tsdf *ts = (lr3d_tsdf *)get_external(env, argv[0]);
    return status_or_undefined(env, which ? lr3d_tsdf_set_depth_culling(ts, on ? 1 : 0)
                                          : lr3d_tsdf_set_counting(ts, on ? 1 : 0));
}
static napi_value tsdf_set_counting(napi_env env, napi_callback_info info) { return tsdf_set_flag(env, info, 0); }
static napi_value tsdf_set_depth_culling(napi_env env, napi_callback_info info) { return tsdf_set_flag(env, info, 1); }

/* tsdfUpdateCount(volume) -> number; tsdfPassCounts(volume) -> { updates, pairs, bricksRead,
 * voxelsWritten } (the counting passes' totals) */
static napi_value tsdf_update_count(napi_env env, napi_callback_info info) {
    ARGS(1);
    int64_t n = 0;
    const int32_t rc = lr3d_tsdf_update_count((lr3d_tsdf *)get_external(env, argv[0]), &n);
    if (rc != LR3D_OK) return throw_status(env, rc);
    napi_value out;
    napi_create_int64(env, n, &out);
    return out;
}

static napi_value tsdf_pass_counts(napi_env env, napi_callback_info info) {
    ARGS(1);
    int64_t c[4] = {0, 0, 0, 0};
    const int32_t rc = lr3d_tsdf_pass_counts((lr3d_tsdf *)get_external(env, argv[0]), c);
    if (rc != LR3D_OK) return throw_status(env, rc);
    static const char *const names[4] = {"updates", "pairs", "bricksRead", "voxelsWritten"};
    napi_value out;
    napi_create_object(env, &out);
    for (int i = 0; i < 4; i++) {
        napi_value v;
        napi_create_int64(env, c[i], &v);
        napi_set_named_property(env, out, names[i], v);
    }
    return out;
}

/* tsdfWrite(volume, tsdf: Float32Array, weight: Float32Array, bricks) (bricks*512 floats each) */
static napi_value tsdf_write(napi_env env, napi_callback_info info) {
    ARGS(4);
    double bricks = 0;
    napi_get_value_double(env, argv[3], &bricks);
    size_t na = 0, nb = 0;
    int bad_a = 0, bad_b = 0;
    float *a = (float *)copy_typed(env, argv[1], napi_float32_array, &na, &bad_a);
    float *b = (float *)copy_typed(env, argv[2], napi_float32_array, &nb, &bad_b);
    if (bad_a || bad_b || !a || !b || (double)na != bricks * 512.0 || (double)nb != bricks * 512.0) {
        free(a);
        free(b);
        napi_throw_range_error(env, NULL, "lr3d.node: tsdf and weight must be Float32Arrays of bricks*512");
        return NULL;
    }
    const int32_t rc = lr3d_tsdf_write((lr3d_tsdf *)get_external(env, argv[0]), a, b);
    free(a);
    free(b);
    return status_or_undefined(env, rc);
}

static napi_value tsdf_read(napi_env env, napi_callback_info info) {
    ARGS(2);
    job_t *j = new_job(OP_READ_TSDF);
    j->ts = (lr3d_tsdf *)get_external(env, argv[0]);
    int32_t bricks = 0;
    napi_get_value_int32(env, argv[1], &bricks);
    j->bytes_a = j->bytes_b = (size_t)bricks * 512u * 4u;
    j->out_a = malloc(j->bytes_a + 1);
    j->out_b = malloc(j->bytes_b + 1);
    return queue(env, j);
}

/* nccBackward(stereo, solveOptions, weights | null, pixels) -> {cost, gradDepth, gradNormal, gradPose, loss} */
static napi_value ncc_backward(napi_env env, napi_callback_info info) {
    ARGS(4);
    job_t *j = new_job(OP_NCC_BACKWARD);
    j->st = (lr3d_stereo *)get_external(env, argv[0]);
    j->n = 1;
    j->descs = (lr3d_solve_desc *)calloc(1, sizeof(lr3d_solve_desc));
    j->priors = (float **)calloc(1, sizeof(float *));
    const int ok = read_solve(env, argv[1], &j->descs[0], &j->priors[0]);
    if (ok <= 0) {
        job_free(j);
        if (ok < 0) napi_throw_range_error(env, NULL, "lr3d.node: prior must be a Float32Array of width*height");
        else napi_throw_type_error(env, NULL, "lr3d.node: solve options need ref and sources");
        return NULL;
    }
    int32_t px = 0;
    napi_get_value_int32(env, argv[3], &px);
    napi_valuetype t = napi_undefined;
    napi_typeof(env, argv[2], &t);
    if (t == napi_object) {
        size_t count = 0;
        int bad = 0;
        j->weights = (float *)copy_typed(env, argv[2], napi_float32_array, &count, &bad);
        if (!j->weights || count != (size_t)px) {
            job_free(j);
            napi_throw_range_error(env, NULL, "lr3d.node: weights must be a Float32Array of width*height");
            return NULL;
        }
    }
    j->bytes_a = j->bytes_b = (size_t)px * 4u;
    j->bytes_c = (size_t)px * 12u;
    j->bytes_d = (size_t)LR3D_MAX_SOURCES * 12u * 8u;
    j->out_a = malloc(j->bytes_a + 1);
    j->out_b = malloc(j->bytes_b + 1);
    j->out_c = malloc(j->bytes_c + 1);
    j->out_d = malloc(j->bytes_d);
    return queue(env, j);
}

/* nccRefineStep(stereo, [solveOptions], stepDepth, stepNormal) -> {gradPose, lossBefore, lossAfter} */
static napi_value ncc_refine_step(napi_env env, napi_callback_info info) {
    ARGS(4);
    job_t *j = new_job(OP_NCC_REFINE);
    j->st = (lr3d_stereo *)get_external(env, argv[0]);
    uint32_t n = 0;
    napi_get_array_length(env, argv[1], &n);
    j->n = (int32_t)n;
    j->descs = (lr3d_solve_desc *)calloc(n ? n : 1, sizeof(lr3d_solve_desc));
    j->priors = (float **)calloc(n ? n : 1, sizeof(float *));
    for (uint32_t i = 0; i < n; i++) {
        napi_value o;
        napi_get_element(env, argv[1], i, &o);
        const int ok = read_solve(env, o, &j->descs[i], &j->priors[i]);
        if (ok <= 0) {
            job_free(j);
            if (ok < 0) napi_throw_range_error(env, NULL, "lr3d.node: prior must be a Float32Array of width*height");
            else napi_throw_type_error(env, NULL, "lr3d.node: solve options need ref and sources");
            return NULL;
        }
    }
    double sd = 0, sn = 0;
    napi_get_value_double(env, argv[2], &sd);
    napi_get_value_double(env, argv[3], &sn);
    j->step_depth = (float)sd;
    j->step_normal = (float)sn;
    j->bytes_a = (size_t)(n ? n : 1) * LR3D_MAX_SOURCES * 12u * 8u;
    j->bytes_b = j->bytes_c = (size_t)(n ? n : 1) * 8u;
    j->out_a = malloc(j->bytes_a);
    j->out_b = malloc(j->bytes_b);
    j->out_c = malloc(j->bytes_c);
    return queue(env, j);
}

/* stereoSetPose(stereo, layer, R[9], t[3]) */
static napi_value stereo_set_pose(napi_env env, napi_callback_info info) {
    ARGS(4);
    lr3d_stereo *st = (lr3d_stereo *)get_external(env, argv[0]);
    int32_t layer = 0;
    napi_get_value_int32(env, argv[1], &layer);
    double R[9], t[3];
    for (uint32_t i = 0; i < 9; i++) {
        napi_value e;
        napi_get_element(env, argv[2], i, &e);
        napi_get_value_double(env, e, &R[i]);
    }
    for (uint32_t i = 0; i < 3; i++) {
        napi_value e;
        napi_get_element(env, argv[3], i, &e);
        napi_get_value_double(env, e, &t[i]);
    }
    int32_t rc = lr3d_stereo_set_pose(st, layer, R, t);
    if (rc != LR3D_OK) return throw_status(env, rc);
    return undefined(env);
}

/* tsdfBackward(tsdf, stereo, layers, map, views, voxels: Uint32Array, targets: Float32Array)
 *   -> {tsdf, gradDepth, gradPose, loss} */
static napi_value tsdf_backward(napi_env env, napi_callback_info info) {
    ARGS(7);
    job_t *j = new_job(OP_TSDF_BACKWARD);
    j->ts = (lr3d_tsdf *)get_external(env, argv[0]);
    j->st = (lr3d_stereo *)get_external(env, argv[1]);
    uint32_t n = 0;
    napi_get_array_length(env, argv[2], &n);
    napi_get_value_int32(env, argv[3], &j->map);
    j->n = (int32_t)n;
    j->layers = (int32_t *)calloc(n ? n : 1, sizeof(int32_t));
    j->iviews = (lr3d_integration_view *)calloc(n ? n : 1, sizeof(lr3d_integration_view));
    j->priors = (float **)calloc(n ? n : 1, sizeof(float *));
    size_t px = 0;
    for (uint32_t i = 0; i < n; i++) {
        napi_value e, o;
        napi_get_element(env, argv[2], i, &e);
        napi_get_value_int32(env, e, &j->layers[i]);
        napi_get_element(env, argv[4], i, &o);
        const int ok = read_iview(env, o, &j->iviews[i], &j->priors[i]);
        if (ok <= 0) {
            job_free(j);
            if (ok < 0) napi_throw_range_error(env, NULL, "lr3d.node: prior must be a Float32Array of width*height");
            else napi_throw_type_error(env, NULL, "lr3d.node: an integration view needs intrinsics, R[9], t[3]");
            return NULL;
        }
        px += (size_t)j->iviews[i].width * j->iviews[i].height;
    }
    size_t nv = 0, nt = 0;
    int bad_v = 0, bad_t = 0;
    j->voxels = (uint32_t *)copy_typed(env, argv[5], napi_uint32_array, &nv, &bad_v);
    j->targets = (float *)copy_typed(env, argv[6], napi_float32_array, &nt, &bad_t);
    if (bad_v || bad_t) {
        job_free(j);
        napi_throw_type_error(env, NULL, "lr3d.node: voxels must be a Uint32Array and targets a Float32Array");
        return NULL;
    }
    if (nv != nt) {
        job_free(j);
        napi_throw_range_error(env, NULL, "lr3d.node: voxels and targets differ in length");
        return NULL;
    }
    j->n_voxels = (int64_t)nv;
    j->bytes_a = nv * 4u;
    j->bytes_b = px * 4u;
    j->bytes_c = (size_t)(n ? n : 1) * 12u * 8u;
    j->out_a = malloc(j->bytes_a + 1);
    j->out_b = malloc(j->bytes_b + 1);
    j->out_c = malloc(j->bytes_c);
    return queue(env, j);
}

/* tsdfExtractSurface(tsdf, minWeight, fraction, minFaces) -> {positions, normals, indices};
 * fraction < 0 skips the small-component removal */
static napi_value tsdf_extract_surface(napi_env env, napi_callback_info info) {
    ARGS(4);
    job_t *j = new_job(OP_MESH);
    j->ts = (lr3d_tsdf *)get_external(env, argv[0]);
    napi_get_value_double(env, argv[1], &j->min_weight);
    napi_get_value_double(env, argv[2], &j->fraction);
    napi_get_value_int32(env, argv[3], &j->min_faces);
    return queue(env, j);
}

/* stereoFuseDepthMaps(stereo, layers, thumbs [{width, height, rgb}], cell, cellBudget, map, minSupport)
 * -> {positions, normals, colors, support, cell} */
static napi_value stereo_fuse_depth_maps(napi_env env, napi_callback_info info) {
    ARGS(7);
    job_t *j = new_job(OP_FUSE);
    j->st = (lr3d_stereo *)get_external(env, argv[0]);
    uint32_t n = 0, nt = 0;
    napi_get_array_length(env, argv[1], &n);
    napi_get_array_length(env, argv[2], &nt);
    if (nt != n) {
        job_free(j);
        napi_throw_range_error(env, NULL, "lr3d.node: one thumbnail per layer");
        return NULL;
    }
    j->n = (int32_t)n;
    j->layers = (int32_t *)calloc(n ? n : 1, sizeof(int32_t));
    j->thumbs = (lr3d_thumb *)calloc(n ? n : 1, sizeof(lr3d_thumb));
    j->rgbs = (uint8_t **)calloc(n ? n : 1, sizeof(uint8_t *));
    for (uint32_t i = 0; i < n; i++) {
        napi_value e, t, rgb;
        napi_get_element(env, argv[1], i, &e);
        napi_get_value_int32(env, e, &j->layers[i]);
        napi_get_element(env, argv[2], i, &t);
        j->thumbs[i].width = (int32_t)get_num(env, t, "width", 0);
        j->thumbs[i].height = (int32_t)get_num(env, t, "height", 0);
        size_t count = 0;
        int bad = 0;
        if (napi_get_named_property(env, t, "rgb", &rgb) == napi_ok)
            j->rgbs[i] = (uint8_t *)copy_typed(env, rgb, napi_uint8_array, &count, &bad);
        if (!j->rgbs[i] || count != (size_t)j->thumbs[i].width * j->thumbs[i].height * 3u) {
            job_free(j);
            napi_throw_type_error(env, NULL, "lr3d.node: a thumbnail needs width, height and rgb (width*height*3 bytes)");
            return NULL;
        }
        j->thumbs[i].rgb = j->rgbs[i];
    }
    double budget = 0;
    napi_get_value_double(env, argv[3], &j->cell);
    napi_get_value_double(env, argv[4], &budget);
    j->cell_budget = (int64_t)budget;
    napi_get_value_int32(env, argv[5], &j->map);
    napi_get_value_int32(env, argv[6], &j->min_support);
    return queue(env, j);
}

static napi_value last_error(napi_env env, napi_callback_info info) {
    (void)info;
    napi_value s;
    napi_create_string_utf8(env, lr3d_last_error(), NAPI_AUTO_LENGTH, &s);
    return s;
}

static napi_value init(napi_env env, napi_value exports) {
    const struct {
        const char *name;
        napi_callback fn;
    } fns[] = {
        {"ctxCreate", ctx_create},           {"ctxDestroy", ctx_destroy},
        {"stereoCreate", stereo_create},     {"stereoDestroy", stereo_destroy},
        {"stereoSetImage", stereo_set_image}, {"stereoSolve", stereo_solve},
        {"stereoFilter", stereo_filter},     {"stereoReadDepth", stereo_read_depth},
        {"stereoReadFiltered", stereo_read_filtered}, {"stereoReadHyp", stereo_read_hyp},
        {"stereoSetCoarse", stereo_set_coarse}, {"tsdfCreate", tsdf_create},
        {"tsdfDestroy", tsdf_destroy},       {"tsdfIntegrate", tsdf_integrate},
        {"tsdfRead", tsdf_read},             {"lastError", last_error},
        {"nccBackward", ncc_backward},       {"nccRefineStep", ncc_refine_step},
        {"stereoSetPose", stereo_set_pose},  {"tsdfBackward", tsdf_backward},
        {"tsdfExtractSurface", tsdf_extract_surface}, {"stereoFuseDepthMaps", stereo_fuse_depth_maps},
        {"stereoSolveNoalloc", stereo_solve_noalloc}, {"stereoAutotune", stereo_autotune},
        {"stereoGetTile", stereo_get_tile},  {"stereoSetTile", stereo_set_tile},
        {"stereoWriteDepth", stereo_write_depth}, {"stereoWriteHyp", stereo_write_hyp},
        {"tsdfAutotune", tsdf_autotune},     {"tsdfIntegrateHost", tsdf_integrate_host},
        {"tsdfGetBricksPerBlock", tsdf_get_bricks_per_block}, {"tsdfSetBricksPerBlock", tsdf_set_bricks_per_block},
        {"tsdfGetXcdChunk", tsdf_get_xcd_chunk}, {"tsdfSetXcdChunk", tsdf_set_xcd_chunk},
        {"tsdfSetCounting", tsdf_set_counting}, {"tsdfSetDepthCulling", tsdf_set_depth_culling},
        {"tsdfUpdateCount", tsdf_update_count}, {"tsdfPassCounts", tsdf_pass_counts},
        {"tsdfWrite", tsdf_write},           {"ctxSynchronize", ctx_synchronize},
        {"ctxSetProfiling", ctx_set_profiling}, {"ctxKernelStats", ctx_kernel_stats},
    };
    for (size_t i = 0; i < sizeof(fns) / sizeof(fns[0]); i++) {
        napi_value f;
        NAPI_CALL(env, napi_create_function(env, fns[i].name, NAPI_AUTO_LENGTH, fns[i].fn, NULL, &f));
        NAPI_CALL(env, napi_set_named_property(env, exports, fns[i].name, f));
    }
    return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)
