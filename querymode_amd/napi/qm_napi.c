/*
 * qm_napi.c — Node-API addon over the C-ABI in include/qmgpu.h.
 *
 * This is the binding a TypeScript host loads to reach the device path: the
 * reference's WASM engine is reached the same way through exported functions
 * over linear memory (src/wasm-engine.ts:21-188), and its native build wraps
 * opaque handles for JS (wasm/src/nodejs.zig:101-742). Handles cross as
 * napi_external values; a handle closed explicitly is not closed again by its
 * finaliser, a handle a pending batch uses closes when the batch settles, and
 * a context shuts down only after its tables, plans and results. Failures throw an Error whose `code` is the QueryModeError code
 * (src/errors.ts:7-32). Results cross as QMCB ArrayBuffers (src/columnar.ts).
 *
 * Exports: init, shutdown, lastError, deviceInfo, tableCreate, tableDrop,
 * tableRows, planCompile, planClose, exec, batchExec, batchSubmit (a Promise of
 * the results; the batch runs while the event loop goes on), planDeviceBytes,
 * resultRows, resultQmcb, resultQmcbRange, resultStats, resultClose.
 */
#define NAPI_VERSION 6
#include <node_api.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/qmgpu.h"

enum { H_CTX = 1, H_TABLE, H_PLAN, H_RESULT };

typedef struct handle {
  int kind;
  void *p;             /* NULL once closed */
  int busy;            /* batches in flight that use it; for a context also its open tables, plans, results */
  int close_pending;   /* closed by the caller while busy: closed when busy drops to 0 */
  int orphaned;        /* finalised by JS while busy: the struct is freed when busy drops to 0 */
  struct handle *parent; /* the context a table, plan or result belongs to (holds one of its busy counts) */
} handle;

static const char *code_name(int rc) {
  switch (rc) {
    case QM_ERR_TABLE_NOT_FOUND: return "TABLE_NOT_FOUND";
    case QM_ERR_COLUMN_NOT_FOUND: return "COLUMN_NOT_FOUND";
    case QM_ERR_INVALID_FORMAT: return "INVALID_FORMAT";
    case QM_ERR_SCHEMA_MISMATCH: return "SCHEMA_MISMATCH";
    case QM_ERR_INVALID_FILTER: return "INVALID_FILTER";
    case QM_ERR_INVALID_AGGREGATE: return "INVALID_AGGREGATE";
    case QM_ERR_MEMORY_EXCEEDED: return "MEMORY_EXCEEDED";
    case QM_ERR_UNSUPPORTED: return "UNSUPPORTED";
    case QM_ERR_INVALID_ARG: return "INVALID_ARGUMENT";
    case QM_ERR_OVERFLOW: return "OVERFLOW";
    default: return "QUERY_FAILED";
  }
}

/* Throw Error(message) with .code; returns NULL for the caller to return. */
static napi_value throw_code(napi_env env, const char *code, const char *msg) {
  napi_value ecode, emsg, err;
  napi_create_string_utf8(env, code, NAPI_AUTO_LENGTH, &ecode);
  napi_create_string_utf8(env, msg, NAPI_AUTO_LENGTH, &emsg);
  napi_create_error(env, ecode, emsg, &err);
  napi_set_named_property(env, err, "code", ecode);
  napi_throw(env, err);
  return NULL;
}

static napi_value throw_rc(napi_env env, int rc) {
  char buf[1024];
  qm_last_error(NULL, buf, sizeof buf);
  return throw_code(env, code_name(rc), buf[0] ? buf : "device call failed");
}

#define CHECK_NAPI(call)                                                     \
  do {                                                                       \
    if ((call) != napi_ok) return throw_code(env, "INVALID_ARGUMENT", #call); \
  } while (0)

static void close_now(handle *h);

/* Drop one busy count; a close asked for meanwhile happens at 0, and a
 * struct JS has already finalised is freed then. */
static void unref(handle *h) {
  if (--h->busy > 0) return;
  if (h->close_pending) close_now(h);
  if (h->orphaned) free(h);
}

/* Close what a handle holds (the library object is gone afterwards), then
 * release the context it belonged to: a context closes only after its
 * tables, plans and results, whose close paths use it. */
static void close_now(handle *h) {
  if (!h->p) return;
  switch (h->kind) {
    case H_TABLE: qm_table_drop((qm_table *)h->p); break;
    case H_PLAN: qm_plan_close((qm_plan *)h->p); break;
    case H_RESULT: qm_result_close((qm_result *)h->p); break;
    case H_CTX: qm_shutdown((qm_ctx *)h->p); break;
  }
  h->p = NULL;
  if (h->parent) {
    handle *par = h->parent;
    h->parent = NULL;
    unref(par);
  }
}

/* A handle a running batch (or, for a context, a live table, plan or result)
 * still uses is closed when the last user lets go. */
static void close_or_defer(handle *h) {
  if (h->busy > 0) h->close_pending = 1;
  else close_now(h);
}

static void finalize_handle(napi_env env, void *data, void *hint) {
  (void)env;
  (void)hint;
  handle *h = (handle *)data;
  if (h->kind != H_CTX) {
    close_now(h); /* a batch holds references to what it uses: never busy here */
    free(h);
  } else if (h->busy > 0) {
    h->orphaned = 1; /* a context is shut down explicitly only */
  } else {
    free(h);
  }
}

/* Wrap a library object; `parent` (a context handle, or NULL) stays open
 * while the new handle is. On failure the object is closed. */
static napi_value wrap(napi_env env, int kind, void *p, handle *parent) {
  handle *h = (handle *)calloc(1, sizeof *h);
  napi_value v;
  if (h) {
    h->kind = kind;
    h->p = p;
    if (napi_create_external(env, h, finalize_handle, NULL, &v) == napi_ok) {
      if (parent) {
        h->parent = parent;
        parent->busy++;
      }
      return v;
    }
    h->p = NULL;
    free(h);
  }
  switch (kind) {
    case H_TABLE: qm_table_drop((qm_table *)p); break;
    case H_PLAN: qm_plan_close((qm_plan *)p); break;
    case H_RESULT: qm_result_close((qm_result *)p); break;
    case H_CTX: qm_shutdown((qm_ctx *)p); break;
  }
  return throw_code(env, "MEMORY_EXCEEDED", "could not wrap a handle");
}

/* The open handle of `kind` in v, or NULL. `use` also rejects a table or plan
 * whose context the caller has shut down: no new work starts on a closing
 * context, while closing them, and reading results, stays allowed. */
static handle *unwrap_as(napi_env env, napi_value v, int kind, int use) {
  napi_valuetype t;
  void *data = NULL;
  if (napi_typeof(env, v, &t) != napi_ok || t != napi_external) return NULL;
  if (napi_get_value_external(env, v, &data) != napi_ok || !data) return NULL;
  handle *h = (handle *)data;
  if (h->kind != kind || !h->p || h->close_pending) return NULL;
  if (use && kind != H_RESULT && h->parent && h->parent->close_pending) return NULL;
  return h;
}
static handle *unwrap(napi_env env, napi_value v, int kind) { return unwrap_as(env, v, kind, 1); }

static int get_args(napi_env env, napi_callback_info info, size_t want, napi_value *argv) {
  size_t argc = want;
  if (napi_get_cb_info(env, info, &argc, argv, NULL, NULL) != napi_ok) return 0;
  for (size_t i = argc; i < want; ++i) napi_get_undefined(env, &argv[i]);
  return 1;
}

static int to_u64(napi_env env, napi_value v, uint64_t *out) {
  napi_valuetype t;
  if (napi_typeof(env, v, &t) != napi_ok) return 0;
  if (t == napi_undefined) { *out = 0; return 1; }
  if (t == napi_bigint) {
    bool lossless;
    return napi_get_value_bigint_uint64(env, v, out, &lossless) == napi_ok;
  }
  double d;
  if (napi_get_value_double(env, v, &d) != napi_ok || d < 0) return 0;
  *out = (uint64_t)d;
  return 1;
}

/* bytes of a TypedArray / ArrayBuffer / Buffer argument */
static int get_bytes(napi_env env, napi_value v, void **data, size_t *len) {
  bool is;
  if (napi_is_typedarray(env, v, &is) == napi_ok && is) {
    napi_typedarray_type tt;
    size_t n, off;
    napi_value ab;
    void *base;
    if (napi_get_typedarray_info(env, v, &tt, &n, &base, &ab, &off) != napi_ok) return 0;
    size_t esz = 1;
    switch (tt) {
      case napi_int16_array: case napi_uint16_array: esz = 2; break;
      case napi_int32_array: case napi_uint32_array: case napi_float32_array: esz = 4; break;
      case napi_float64_array: case napi_bigint64_array: case napi_biguint64_array: esz = 8; break;
      default: esz = 1;
    }
    *data = base;
    *len = n * esz;
    return 1;
  }
  if (napi_is_arraybuffer(env, v, &is) == napi_ok && is)
    return napi_get_arraybuffer_info(env, v, data, len) == napi_ok;
  return 0;
}

static napi_value js_init(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return throw_code(env, "INVALID_ARGUMENT", "bad arguments");
  int32_t dev = 0;
  napi_valuetype t;
  napi_typeof(env, argv[0], &t);
  if (t != napi_undefined) CHECK_NAPI(napi_get_value_int32(env, argv[0], &dev));
  uint64_t budget = 0;
  if (!to_u64(env, argv[1], &budget)) return throw_code(env, "INVALID_ARGUMENT", "budget must be a number");
  qm_ctx *ctx = NULL;
  int rc = qm_init(dev, budget, &ctx);
  if (rc != QM_OK) return throw_rc(env, rc);
  return wrap(env, H_CTX, ctx, NULL);
}

static napi_value js_shutdown(napi_env env, napi_callback_info info) {
  napi_value argv[1], undef;
  get_args(env, info, 1, argv);
  handle *h = unwrap(env, argv[0], H_CTX);
  if (h) close_or_defer(h);
  napi_get_undefined(env, &undef);
  return undef;
}

static napi_value js_last_error(napi_env env, napi_callback_info info) {
  (void)info;
  char buf[4096];
  napi_value s;
  qm_last_error(NULL, buf, sizeof buf);
  napi_create_string_utf8(env, buf, NAPI_AUTO_LENGTH, &s);
  return s;
}

static napi_value js_device_info(napi_env env, napi_callback_info info) {
  napi_value argv[1], o, v;
  get_args(env, info, 1, argv);
  handle *h = unwrap(env, argv[0], H_CTX);
  if (!h) return throw_code(env, "INVALID_ARGUMENT", "expected a context");
  char name[128];
  uint32_t cus = 0;
  uint64_t budget = 0, used = 0;
  int rc = qm_device_info((qm_ctx *)h->p, name, sizeof name, &cus, &budget, &used);
  if (rc != QM_OK) return throw_rc(env, rc);
  napi_create_object(env, &o);
  napi_create_string_utf8(env, name, NAPI_AUTO_LENGTH, &v);
  napi_set_named_property(env, o, "arch", v);
  napi_create_uint32(env, cus, &v);
  napi_set_named_property(env, o, "cus", v);
  napi_create_double(env, (double)budget, &v);
  napi_set_named_property(env, o, "budget", v);
  napi_create_double(env, (double)used, &v);
  napi_set_named_property(env, o, "used", v);
  return o;
}

/* tableCreate(ctx, name, nrows, [{name, dtype, data, nulls?}, ...]) */
static napi_value js_table_create(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  if (!get_args(env, info, 4, argv)) return throw_code(env, "INVALID_ARGUMENT", "bad arguments");
  handle *h = unwrap(env, argv[0], H_CTX);
  if (!h) return throw_code(env, "INVALID_ARGUMENT", "expected a context");
  char tname[256];
  size_t tl;
  CHECK_NAPI(napi_get_value_string_utf8(env, argv[1], tname, sizeof tname, &tl));
  uint64_t nrows = 0;
  if (!to_u64(env, argv[2], &nrows)) return throw_code(env, "INVALID_ARGUMENT", "nrows must be a number");
  uint32_t ncols = 0;
  CHECK_NAPI(napi_get_array_length(env, argv[3], &ncols));
  qm_column_desc *cols = (qm_column_desc *)calloc(ncols ? ncols : 1, sizeof *cols);
  char (*names)[256] = (char (*)[256])calloc(ncols ? ncols : 1, 256);
  if (!cols || !names) {
    free(cols);
    free(names);
    return throw_code(env, "MEMORY_EXCEEDED", "out of host memory");
  }
  napi_value result = NULL;
  for (uint32_t i = 0; i < ncols; ++i) {
    napi_value c, v;
    size_t nl, len;
    void *data;
    int32_t dtype;
    bool has;
    if (napi_get_element(env, argv[3], i, &c) != napi_ok ||
        napi_get_named_property(env, c, "name", &v) != napi_ok ||
        napi_get_value_string_utf8(env, v, names[i], 256, &nl) != napi_ok ||
        napi_get_named_property(env, c, "dtype", &v) != napi_ok ||
        napi_get_value_int32(env, v, &dtype) != napi_ok ||
        napi_get_named_property(env, c, "data", &v) != napi_ok || !get_bytes(env, v, &data, &len)) {
      throw_code(env, "INVALID_ARGUMENT", "column needs {name, dtype, data: TypedArray}");
      goto done;
    }
    cols[i].name = names[i];
    cols[i].dtype = dtype;
    cols[i].data = data;
    if (napi_has_named_property(env, c, "nulls", &has) == napi_ok && has) {
      napi_valuetype t;
      napi_get_named_property(env, c, "nulls", &v);
      napi_typeof(env, v, &t);
      if (t != napi_undefined && t != napi_null) {
        void *nb;
        size_t nbl;
        if (!get_bytes(env, v, &nb, &nbl) || nbl < (nrows + 7) / 8) {
          throw_code(env, "INVALID_ARGUMENT", "nulls must be a Uint8Array of ceil(nrows/8) bytes");
          goto done;
        }
        cols[i].null_bitmap = (const uint8_t *)nb;
      }
    }
  }
  {
    qm_table *t = NULL;
    int rc = qm_table_create((qm_ctx *)h->p, tname, nrows, ncols, cols, &t);
    if (rc != QM_OK) throw_rc(env, rc);
    else result = wrap(env, H_TABLE, t, h);
  }
done:
  free(cols);
  free(names);
  return result;
}

static napi_value close_kind(napi_env env, napi_callback_info info, int kind) {
  napi_value argv[1], undef;
  get_args(env, info, 1, argv);
  handle *h = unwrap_as(env, argv[0], kind, 0);
  if (h) close_or_defer(h);
  napi_get_undefined(env, &undef);
  return undef;
}
static napi_value js_table_drop(napi_env env, napi_callback_info info) { return close_kind(env, info, H_TABLE); }
static napi_value js_plan_close(napi_env env, napi_callback_info info) { return close_kind(env, info, H_PLAN); }
static napi_value js_result_close(napi_env env, napi_callback_info info) { return close_kind(env, info, H_RESULT); }

static napi_value js_table_rows(napi_env env, napi_callback_info info) {
  napi_value argv[1], v;
  get_args(env, info, 1, argv);
  handle *h = unwrap(env, argv[0], H_TABLE);
  if (!h) return throw_code(env, "INVALID_ARGUMENT", "expected a table");
  napi_create_double(env, (double)qm_table_rows((qm_table *)h->p), &v);
  return v;
}

static napi_value js_plan_compile(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  get_args(env, info, 2, argv);
  handle *h = unwrap(env, argv[0], H_CTX);
  void *data;
  size_t len;
  if (!h || !get_bytes(env, argv[1], &data, &len))
    return throw_code(env, "INVALID_ARGUMENT", "planCompile(ctx, Uint8Array)");
  qm_plan *p = NULL;
  int rc = qm_plan_compile((qm_ctx *)h->p, (const uint8_t *)data, len, &p);
  if (rc != QM_OK) return throw_rc(env, rc);
  return wrap(env, H_PLAN, p, h);
}

static napi_value js_exec(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  get_args(env, info, 3, argv);
  handle *p = unwrap(env, argv[0], H_PLAN);
  handle *l = unwrap(env, argv[1], H_TABLE);
  handle *r = unwrap(env, argv[2], H_TABLE);
  if (!p || !l) return throw_code(env, "INVALID_ARGUMENT", "exec(plan, table[, rightTable])");
  qm_result *res = NULL;
  int rc = qm_exec((qm_plan *)p->p, (qm_table *)l->p, r ? (qm_table *)r->p : NULL, &res);
  if (rc != QM_OK) return throw_rc(env, rc);
  return wrap(env, H_RESULT, res, p->parent);
}

/* planDeviceBytes(plan, table[, rightTable]) -> device bytes beyond the resident tables */
static napi_value js_plan_device_bytes(napi_env env, napi_callback_info info) {
  napi_value argv[3], v;
  get_args(env, info, 3, argv);
  handle *p = unwrap(env, argv[0], H_PLAN);
  handle *l = unwrap(env, argv[1], H_TABLE);
  handle *r = unwrap(env, argv[2], H_TABLE);
  if (!p || !l) return throw_code(env, "INVALID_ARGUMENT", "planDeviceBytes(plan, table[, rightTable])");
  uint64_t b = 0;
  int rc = qm_plan_device_bytes((qm_plan *)p->p, (qm_table *)l->p, r ? (qm_table *)r->p : NULL, &b);
  if (rc != QM_OK) return throw_rc(env, rc);
  napi_create_double(env, (double)b, &v);
  return v;
}

/* batchExec(ctx, [plan...], table) -> [result...] */
static napi_value js_batch_exec(napi_env env, napi_callback_info info) {
  napi_value argv[3], arr;
  get_args(env, info, 3, argv);
  handle *c = unwrap(env, argv[0], H_CTX);
  handle *t = unwrap(env, argv[2], H_TABLE);
  uint32_t n = 0;
  if (!c || !t || napi_get_array_length(env, argv[1], &n) != napi_ok)
    return throw_code(env, "INVALID_ARGUMENT", "batchExec(ctx, plans[], table)");
  qm_plan **plans = (qm_plan **)calloc(n ? n : 1, sizeof *plans);
  qm_result **res = (qm_result **)calloc(n ? n : 1, sizeof *res);
  napi_value out = NULL;
  if (!plans || !res) {
    throw_code(env, "MEMORY_EXCEEDED", "out of host memory");
    goto done;
  }
  for (uint32_t i = 0; i < n; ++i) {
    napi_value v;
    handle *ph;
    if (napi_get_element(env, argv[1], i, &v) != napi_ok || !(ph = unwrap(env, v, H_PLAN))) {
      throw_code(env, "INVALID_ARGUMENT", "plans[] must hold plan handles");
      goto done;
    }
    plans[i] = (qm_plan *)ph->p;
  }
  {
    int rc = qm_batch_exec((qm_ctx *)c->p, plans, n, (qm_table *)t->p, res);
    if (rc != QM_OK) {
      throw_rc(env, rc);
      goto done;
    }
  }
  napi_create_array_with_length(env, n, &arr);
  for (uint32_t i = 0; i < n; ++i) napi_set_element(env, arr, i, wrap(env, H_RESULT, res[i], c));
  out = arr;
done:
  free(plans);
  free(res);
  return out;
}

/* batchSubmit(ctx, [plan...], table) -> Promise<[result...]>: qm_batch_submit
 * on the JS thread, qm_batch_wait on a libuv pool thread; the plans array and
 * the table are referenced until the promise settles (a finalised plan would
 * otherwise be closed under the running batch). */
typedef struct batch_work {
  napi_async_work work;
  napi_deferred deferred;
  napi_ref keep_ctx, keep_table;
  napi_ref *keep_plan;          /* each plan element (not only the array) */
  handle *hctx, *htable;
  handle **hplan;
  qm_batch *batch;
  uint32_t n;
  qm_result **res;
  int rc;
  char err[1024];
} batch_work;

static void batch_wait_exec(napi_env env, void *data) {
  (void)env;
  batch_work *w = (batch_work *)data;
  w->rc = qm_batch_wait(w->batch, w->res);
  if (w->rc != QM_OK) qm_last_error(NULL, w->err, sizeof w->err);  /* this pool thread's message */
  qm_batch_close(w->batch);
  w->batch = NULL;
}

static void batch_wait_done(napi_env env, napi_status status, void *data) {
  batch_work *w = (batch_work *)data;
  if (status == napi_ok && w->rc == QM_OK) {
    napi_value arr;
    napi_create_array_with_length(env, w->n, &arr);
    for (uint32_t i = 0; i < w->n; ++i) napi_set_element(env, arr, i, wrap(env, H_RESULT, w->res[i], w->hctx));
    napi_resolve_deferred(env, w->deferred, arr);
  } else {
    napi_value ecode, emsg, err;
    const char *code = w->rc == QM_OK ? "QUERY_FAILED" : code_name(w->rc);
    napi_create_string_utf8(env, code, NAPI_AUTO_LENGTH, &ecode);
    napi_create_string_utf8(env, w->err[0] ? w->err : "batch failed", NAPI_AUTO_LENGTH, &emsg);
    napi_create_error(env, ecode, emsg, &err);
    napi_set_named_property(env, err, "code", ecode);
    napi_reject_deferred(env, w->deferred, err);
  }
  /* the batch is done with its handles: closes the caller asked for meanwhile happen now
     (plans and the table before the context, which closes after them) */
  for (uint32_t i = 0; i < w->n; ++i) {
    unref(w->hplan[i]);
    napi_delete_reference(env, w->keep_plan[i]);
  }
  unref(w->htable);
  unref(w->hctx); /* the results just wrapped hold the context open until they close */
  napi_delete_reference(env, w->keep_table);
  napi_delete_reference(env, w->keep_ctx);
  napi_delete_async_work(env, w->work);
  free(w->keep_plan);
  free(w->hplan);
  free(w->res);
  free(w);
}

static napi_value js_batch_submit(napi_env env, napi_callback_info info) {
  napi_value argv[3], promise, name;
  get_args(env, info, 3, argv);
  handle *c = unwrap(env, argv[0], H_CTX);
  handle *t = unwrap(env, argv[2], H_TABLE);
  uint32_t n = 0;
  if (!c || !t || napi_get_array_length(env, argv[1], &n) != napi_ok || n == 0)
    return throw_code(env, "INVALID_ARGUMENT", "batchSubmit(ctx, plans[], table)");
  qm_plan **plans = (qm_plan **)calloc(n, sizeof *plans);
  batch_work *w = (batch_work *)calloc(1, sizeof *w);
  qm_result **res = (qm_result **)calloc(n, sizeof *res);
  handle **hplan = (handle **)calloc(n, sizeof *hplan);
  napi_value *pv = (napi_value *)calloc(n, sizeof *pv);
  napi_ref *keep_plan = (napi_ref *)calloc(n, sizeof *keep_plan);
  if (!plans || !w || !res || !hplan || !pv || !keep_plan) {
    free(plans); free(w); free(res); free(hplan); free(pv); free(keep_plan);
    return throw_code(env, "MEMORY_EXCEEDED", "out of host memory");
  }
  for (uint32_t i = 0; i < n; ++i) {
    handle *ph;
    if (napi_get_element(env, argv[1], i, &pv[i]) != napi_ok || !(ph = unwrap(env, pv[i], H_PLAN))) {
      free(plans); free(w); free(res); free(hplan); free(pv); free(keep_plan);
      return throw_code(env, "INVALID_ARGUMENT", "plans[] must hold plan handles");
    }
    plans[i] = (qm_plan *)ph->p;
    hplan[i] = ph;
  }
  int rc = qm_batch_submit((qm_ctx *)c->p, plans, n, (qm_table *)t->p, &w->batch);
  free(plans);
  if (rc != QM_OK) {
    free(w); free(res); free(hplan); free(pv); free(keep_plan);
    return throw_rc(env, rc);
  }
  w->n = n;
  w->res = res;
  /* the handles stay open (and referenced, so not finalised) until the batch settles */
  w->hctx = c;
  w->htable = t;
  w->hplan = hplan;
  w->keep_plan = keep_plan;
  c->busy++;
  t->busy++;
  for (uint32_t i = 0; i < n; ++i) {
    hplan[i]->busy++;
    napi_create_reference(env, pv[i], 1, &keep_plan[i]);
  }
  free(pv);
  napi_create_reference(env, argv[0], 1, &w->keep_ctx);
  napi_create_reference(env, argv[2], 1, &w->keep_table);
  napi_create_promise(env, &w->deferred, &promise);
  napi_create_string_utf8(env, "qm_batch", NAPI_AUTO_LENGTH, &name);
  napi_create_async_work(env, NULL, name, batch_wait_exec, batch_wait_done, w, &w->work);
  napi_queue_async_work(env, w->work);
  return promise;
}

static napi_value js_result_rows(napi_env env, napi_callback_info info) {
  napi_value argv[1], v;
  get_args(env, info, 1, argv);
  handle *h = unwrap(env, argv[0], H_RESULT);
  if (!h) return throw_code(env, "INVALID_ARGUMENT", "expected a result");
  uint64_t n = 0;
  int rc = qm_result_rows((qm_result *)h->p, &n);
  if (rc != QM_OK) return throw_rc(env, rc);
  napi_create_double(env, (double)n, &v);
  return v;
}

static void release_qmcb(napi_env env, void *data, void *hint) {
  (void)env;
  (void)hint;
  qm_qmcb_release((const uint8_t *)data);
}

/* QMCB of rows [row0, row0 + nrows) encoded on the device into pinned host
 * memory, handed to JS as an external ArrayBuffer over those bytes (no copy);
 * the buffer goes back to the library when the ArrayBuffer is collected. */
static napi_value qmcb_range(napi_env env, handle *h, uint64_t row0, uint64_t nrows) {
  const uint8_t *p = NULL;
  uint64_t size = 0;
  napi_value ab;
  int rc = qm_result_qmcb_pinned((qm_result *)h->p, row0, nrows, &p, &size);
  if (rc != QM_OK) return throw_rc(env, rc);
  if (napi_create_external_arraybuffer(env, (void *)p, (size_t)size, release_qmcb, NULL, &ab) != napi_ok) {
    qm_qmcb_release(p);
    return throw_code(env, "QUERY_FAILED", "cannot wrap the QMCB buffer");
  }
  return ab;
}

/* resultQmcb(result) -> ArrayBuffer holding the QMCB bytes of the whole result */
static napi_value js_result_qmcb(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  get_args(env, info, 1, argv);
  handle *h = unwrap(env, argv[0], H_RESULT);
  if (!h) return throw_code(env, "INVALID_ARGUMENT", "expected a result");
  return qmcb_range(env, h, 0, UINT64_MAX);
}

/* resultQmcbRange(result, row0, nrows) -> ArrayBuffer of QMCB for those rows (a cursor's batch) */
static napi_value js_result_qmcb_range(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  get_args(env, info, 3, argv);
  handle *h = unwrap(env, argv[0], H_RESULT);
  if (!h) return throw_code(env, "INVALID_ARGUMENT", "expected a result");
  double r0 = 0, n = 0;
  if (napi_get_value_double(env, argv[1], &r0) != napi_ok || napi_get_value_double(env, argv[2], &n) != napi_ok ||
      r0 < 0 || n < 0)
    return throw_code(env, "INVALID_ARGUMENT", "row0 and nrows must be non-negative numbers");
  return qmcb_range(env, h, (uint64_t)r0, (uint64_t)n);
}

static void set_num(napi_env env, napi_value o, const char *k, double x) {
  napi_value v;
  napi_create_double(env, x, &v);
  napi_set_named_property(env, o, k, v);
}

static napi_value js_result_stats(napi_env env, napi_callback_info info) {
  napi_value argv[1], o;
  get_args(env, info, 1, argv);
  handle *h = unwrap(env, argv[0], H_RESULT);
  if (!h) return throw_code(env, "INVALID_ARGUMENT", "expected a result");
  qm_stats st;
  int rc = qm_result_stats((qm_result *)h->p, &st);
  if (rc != QM_OK) return throw_rc(env, rc);
  napi_create_object(env, &o);
  set_num(env, o, "rowsScanned", (double)st.rows_scanned);
  set_num(env, o, "rowsSelected", (double)st.rows_selected);
  set_num(env, o, "groups", (double)st.groups);
  set_num(env, o, "bytesAlgorithmic", (double)st.bytes_algorithmic);
  set_num(env, o, "deviceMs", st.device_ms);
  set_num(env, o, "scanMs", st.scan_ms);
  set_num(env, o, "wallMs", st.wall_ms);
  set_num(env, o, "kernelLaunches", st.kernel_launches);
  set_num(env, o, "retries", st.retries);
  set_num(env, o, "strategy", st.strategy);
  set_num(env, o, "pagesSkipped", st.pages_skipped);
  return o;
}

static napi_value init_module(napi_env env, napi_value exports) {
  napi_property_descriptor props[] = {
    {"init", NULL, js_init, NULL, NULL, NULL, napi_enumerable, NULL},
    {"shutdown", NULL, js_shutdown, NULL, NULL, NULL, napi_enumerable, NULL},
    {"lastError", NULL, js_last_error, NULL, NULL, NULL, napi_enumerable, NULL},
    {"deviceInfo", NULL, js_device_info, NULL, NULL, NULL, napi_enumerable, NULL},
    {"tableCreate", NULL, js_table_create, NULL, NULL, NULL, napi_enumerable, NULL},
    {"tableDrop", NULL, js_table_drop, NULL, NULL, NULL, napi_enumerable, NULL},
    {"tableRows", NULL, js_table_rows, NULL, NULL, NULL, napi_enumerable, NULL},
    {"planCompile", NULL, js_plan_compile, NULL, NULL, NULL, napi_enumerable, NULL},
    {"planClose", NULL, js_plan_close, NULL, NULL, NULL, napi_enumerable, NULL},
    {"exec", NULL, js_exec, NULL, NULL, NULL, napi_enumerable, NULL},
    {"batchExec", NULL, js_batch_exec, NULL, NULL, NULL, napi_enumerable, NULL},
    {"batchSubmit", NULL, js_batch_submit, NULL, NULL, NULL, napi_enumerable, NULL},
    {"planDeviceBytes", NULL, js_plan_device_bytes, NULL, NULL, NULL, napi_enumerable, NULL},
    {"resultRows", NULL, js_result_rows, NULL, NULL, NULL, napi_enumerable, NULL},
    {"resultQmcb", NULL, js_result_qmcb, NULL, NULL, NULL, napi_enumerable, NULL},
    {"resultQmcbRange", NULL, js_result_qmcb_range, NULL, NULL, NULL, napi_enumerable, NULL},
    {"resultStats", NULL, js_result_stats, NULL, NULL, NULL, napi_enumerable, NULL},
    {"resultClose", NULL, js_result_close, NULL, NULL, NULL, napi_enumerable, NULL},
  };
  if (napi_define_properties(env, exports, sizeof props / sizeof props[0], props) != napi_ok) return NULL;
  return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, init_module)
