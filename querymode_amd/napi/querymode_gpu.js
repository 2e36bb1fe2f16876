'use strict';
/*
 * querymode_gpu.js — JavaScript host of the device path over the Node-API
 * addon (qm_napi.node): the GPU counterparts of the reference's
 * QueryExecutor (src/client.ts:1280-1304), pull-based Operator
 * (src/operators.ts:67-90) and DataFrame builder (src/client.ts:176-312), with
 * the binary plan encoding read by qm_plan_compile and a QMCB decoder
 * (src/columnar.ts:438-613). CommonJS and Node 12 syntax.
 *
 * Literal typing follows JavaScript itself: a `number` literal is encoded as
 * an f64, a `bigint` literal as an exact int64 — the distinction the
 * reference's filters make (src/operators.ts:487-547, src/decode.ts:476-556).
 */
const path = require('path');

const addon = require(path.join(__dirname, '..', 'lib', 'qm_napi.node'));

const OPS = { eq: 0, neq: 1, lt: 2, lte: 3, gt: 4, gte: 5, between: 6, not_between: 7, in: 8, not_in: 9,
  is_null: 10, is_not_null: 11, like: 12, not_like: 13 };
const AGG_FNS = { sum: 0, avg: 1, min: 2, max: 3, count: 4, count_distinct: 5, stddev: 6, variance: 7,
  median: 8, percentile: 9 };
const JOIN_TYPES = { inner: 0, left: 1, right: 2, full: 3, cross: 4 };
const PLAN_MAGIC = 0x4C504D51;
const DTYPE = { f64: 0, i64: 1, utf8: 2, bool: 3, f32vec: 4, null: 5, i32: 6, f32: 7, i8: 16, i16: 17 };

function qmError(code, message) {
  const e = new Error(message);
  e.code = code;
  return e;
}

// ---- plan encoding (little-endian; the layout qm_plan_compile reads) ----
class Writer {
  constructor() { this.parts = []; }
  u8(v) { this.parts.push(Buffer.from([v & 0xff])); }
  u16(v) { const b = Buffer.alloc(2); b.writeUInt16LE(v, 0); this.parts.push(b); }
  u32(v) { const b = Buffer.alloc(4); b.writeUInt32LE(v, 0); this.parts.push(b); }
  i64(v) { const b = Buffer.alloc(8); b.writeBigInt64LE(BigInt(v), 0); this.parts.push(b); }
  f64(v) { const b = Buffer.alloc(8); b.writeDoubleLE(v, 0); this.parts.push(b); }
  str(s) { const b = Buffer.from(String(s), 'utf8'); this.u16(b.length); this.parts.push(b); }
  bytes() { return Buffer.concat(this.parts); }
}

function writeLit(w, v) {
  if (typeof v === 'bigint') {
    if (v < -(1n << 63n) || v >= (1n << 63n)) throw qmError('INVALID_FILTER', 'bigint literal outside int64');
    w.u8(2); w.i64(v);
  } else if (typeof v === 'number') {
    w.u8(1); w.f64(v);
  } else {
    throw qmError('UNSUPPORTED', 'only number and bigint literals are on the device path');
  }
}

function writeFilter(w, f) {
  if (!(f.op in OPS)) throw qmError('INVALID_FILTER', 'unknown filter op ' + f.op);
  if (f.op === 'like' || f.op === 'not_like') {
    throw qmError('UNSUPPORTED', 'LIKE needs string columns, which are not on the device path');
  }
  w.str(f.column); w.u8(OPS[f.op]);
  if (f.op === 'is_null' || f.op === 'is_not_null') { w.u8(0); return; }
  if (f.op === 'between' || f.op === 'not_between') {
    if (!Array.isArray(f.value) || f.value.length !== 2) throw qmError('INVALID_FILTER', 'between needs [lo, hi]');
    w.u8(4); writeLit(w, f.value[0]); writeLit(w, f.value[1]); return;
  }
  if (f.op === 'in' || f.op === 'not_in') {
    const vals = Array.from(f.value);
    w.u8(3); w.u32(vals.length); vals.forEach((v) => writeLit(w, v)); return;
  }
  writeLit(w, f.value);
}

function writeFilters(w, fs) { fs = fs || []; w.u16(fs.length); fs.forEach((f) => writeFilter(w, f)); }

function aggName(a) { return a.alias !== undefined && a.alias !== null ? a.alias : a.fn + '_' + a.column; }

// HAVING tree -> postfix ops (the nodes of querymode_amd/plan.py HAVING_NODES)
const HCMP = { eq: 0, ne: 1, lt: 2, le: 3, gt: 4, ge: 5 };
const HK = { cmp: 1, cmp_col: 2, between: 3, in: 4, isnull: 5, and: 6, or: 7, not: 8 };
function havingOps(e, out) {
  const k = e[0];
  if (!(k in HK)) throw qmError('INVALID_FORMAT', 'unknown HAVING node ' + k);
  const w = new Writer();
  if (k === 'and' || k === 'or') { havingOps(e[1], out); havingOps(e[2], out); w.u8(HK[k]); }
  else if (k === 'not') { havingOps(e[1], out); w.u8(HK[k]); }
  else if (k === 'cmp') { w.u8(HK[k]); w.str(e[2]); w.u8(HCMP[e[1]]); writeLit(w, e[3]); }
  else if (k === 'cmp_col') { w.u8(HK[k]); w.str(e[2]); w.u8(HCMP[e[1]]); w.str(e[3]); }
  else if (k === 'between') { w.u8(HK[k]); w.str(e[1]); w.u8(e[4] ? 1 : 0); writeLit(w, e[2]); writeLit(w, e[3]); }
  else if (k === 'in') {
    w.u8(HK[k]); w.str(e[1]); w.u8(e[3] ? 1 : 0); w.u16(e[2].length);
    e[2].forEach((v) => { if (v === null) w.u8(0); else writeLit(w, v); });
  } else { w.u8(HK[k]); w.str(e[1]); w.u8(e[2] ? 1 : 0); }
  out.push(w.bytes());
}

/** QueryDescriptor (src/client.ts:1254-1277) -> plan bytes. */
function encodePlan(d) {
  const w = new Writer();
  const distinct = d.distinct === undefined ? null : d.distinct;
  const orderBy = d.orderBy || [];
  const v2 = distinct !== null || orderBy.length > 0 || (d.having !== undefined && d.having !== null);
  w.u32(PLAN_MAGIC); w.u16(v2 ? 2 : 1);
  writeFilters(w, d.filters);
  const groups = d.filterGroups || [];
  w.u16(groups.length); groups.forEach((g) => writeFilters(w, g));
  const comp = d.computedColumns || [];   // product expressions: [alias, a, b]
  w.u16(comp.length);
  comp.forEach((c) => { w.str(c[0]); w.u8(1); w.str(c[1]); w.str(c[2]); });
  const gb = d.groupBy || [];
  w.u16(gb.length); gb.forEach((g) => w.str(g));
  const aggs = d.aggregates || [];
  w.u16(aggs.length);
  aggs.forEach((a) => {
    if (!(a.fn in AGG_FNS)) throw qmError('INVALID_AGGREGATE', 'unknown aggregate ' + a.fn);
    w.u8(AGG_FNS[a.fn]); w.str(a.column); w.str(aggName(a));
  });
  if (d.join) {
    w.u8(1); w.str(d.join.leftKey); w.str(d.join.rightKey); w.u8(JOIN_TYPES[d.join.type || 'inner']);
    writeFilters(w, d.join.right.filters); const rg = d.join.right.filterGroups || [];
    w.u16(rg.length); rg.forEach((g) => writeFilters(w, g));
  } else {
    w.u8(0);
  }
  if (d.sortColumn) { w.u8(1); w.str(d.sortColumn); w.u8(d.sortDirection === 'desc' ? 1 : 0); } else { w.u8(0); }
  w.i64(d.limit === undefined || d.limit === null ? -1 : d.limit);
  w.i64(d.offset || 0);
  const proj = d.projections || [];
  w.u16(proj.length); proj.forEach((p) => w.str(p));
  if (v2) {
    // SQL post-stages: DISTINCT columns, multi-column ORDER BY, HAVING program
    if (distinct !== null) { w.u8(1); w.u16(distinct.length); distinct.forEach((c) => w.str(c)); } else { w.u8(0); }
    w.u16(orderBy.length); orderBy.forEach((o) => { w.str(o[0]); w.u8(o[1] === 'desc' ? 1 : 0); });
    const ops = [];
    if (d.having !== undefined && d.having !== null) havingOps(d.having, ops);
    w.u16(ops.length); ops.forEach((b) => w.parts.push(b));
  }
  return w.bytes();
}

// ---- QMCB (src/columnar.ts:438-613) ----
function decodeQMCB(ab) {
  const buf = Buffer.from(ab);
  if (buf.length < 10 || buf.readUInt32LE(0) !== 0x42434D51) return null;
  const rows = buf.readUInt32LE(4);
  const ncols = buf.readUInt16LE(8);
  let pos = 10;
  const descs = [];
  for (let i = 0; i < ncols; i++) {
    const nl = buf.readUInt16LE(pos); pos += 2;
    const name = buf.toString('utf8', pos, pos + nl); pos += nl;
    descs.push({ name, dtype: buf[pos], hasNulls: buf[pos + 1] !== 0 }); pos += 2;
  }
  const columns = [];
  for (const d of descs) {
    let nullBitmap = null;
    if (d.hasNulls) { const k = (rows + 7) >> 3; nullBitmap = Uint8Array.from(buf.subarray(pos, pos + k)); pos += k; }
    let data;
    const take = (n) => { const b = Buffer.from(buf.subarray(pos, pos + n)); pos += n; return b; };
    if (d.dtype === DTYPE.f64) { const b = take(rows * 8); data = new Float64Array(b.buffer, b.byteOffset, rows); }
    else if (d.dtype === DTYPE.i64) { const b = take(rows * 8); data = new BigInt64Array(b.buffer, b.byteOffset, rows); }
    else if (d.dtype === DTYPE.i32) { const b = take(rows * 4); data = new Int32Array(b.buffer, b.byteOffset, rows); }
    else if (d.dtype === DTYPE.f32) { const b = take(rows * 4); data = new Float32Array(b.buffer, b.byteOffset, rows); }
    else if (d.dtype === DTYPE.bool) { const b = take((rows + 7) >> 3); data = Array.from({ length: rows }, (_, i) => ((b[i >> 3] >> (i & 7)) & 1) === 1); }
    else throw qmError('INVALID_FORMAT', 'QMCB dtype ' + d.dtype + ' not produced by the device path');
    columns.push({ name: d.name, dtype: d.dtype, data, nullBitmap, rowCount: rows });
  }
  return { rowCount: rows, columns };
}

/** columnarBatchToRows: NULL -> null, I64 -> bigint (src/columnar.ts:1056-1121). */
function qmcbToRows(batch) {
  const rows = [];
  for (let i = 0; i < batch.rowCount; i++) rows.push({});
  for (const c of batch.columns) {
    for (let i = 0; i < batch.rowCount; i++) {
      const isNull = c.nullBitmap && ((c.nullBitmap[i >> 3] >> (i & 7)) & 1) === 1;
      rows[i][c.name] = isNull ? null : c.data[i];
    }
  }
  return rows;
}

// ---- executor / operator / builder ----
const NP = { Float64Array: DTYPE.f64, BigInt64Array: DTYPE.i64, Int32Array: DTYPE.i32, Float32Array: DTYPE.f32,
  Int8Array: DTYPE.i8, Int16Array: DTYPE.i16, Uint8Array: DTYPE.bool };

class GpuExecutor {
  constructor(device, hbmBudgetBytes) {
    this.ctx = addon.init(device || 0, hbmBudgetBytes || 0);
    this.tables = new Map();
  }

  /** Upload decoded columns once: {name: TypedArray}, optional nulls {name: Uint8Array bitmap}. */
  registerTable(name, columns, nulls) {
    const names = Object.keys(columns);
    const nrows = names.length ? columns[names[0]].length : 0;
    const cols = names.map((n) => {
      const a = columns[n];
      const dtype = NP[a.constructor.name];
      if (dtype === undefined) throw qmError('UNSUPPORTED', 'column ' + n + ': ' + a.constructor.name);
      if (a.length !== nrows) throw qmError('SCHEMA_MISMATCH', 'column ' + n + ' has ' + a.length + ' rows');
      return { name: n, dtype, data: a, nulls: nulls && nulls[n] ? nulls[n] : undefined };
    });
    const old = this.tables.get(name);
    if (old) addon.tableDrop(old);
    const t = addon.tableCreate(this.ctx, name, nrows, cols);
    this.tables.set(name, t);
    return t;
  }

  _run(desc) {
    const left = this.tables.get(desc.table);
    if (!left) throw qmError('TABLE_NOT_FOUND', 'table ' + desc.table + ' is not registered');
    let right = null;
    if (desc.join) {
      right = this.tables.get(desc.join.right.table);
      if (!right) throw qmError('TABLE_NOT_FOUND', 'table ' + desc.join.right.table + ' is not registered');
    }
    const plan = addon.planCompile(this.ctx, encodePlan(desc));
    try {
      const res = addon.exec(plan, left, right);
      try {
        return { qmcb: addon.resultQmcb(res), stats: addon.resultStats(res) };
      } finally { addon.resultClose(res); }
    } finally { addon.planClose(plan); }
  }

  /**
   * QueryExecutor.cursor (src/client.ts:1300): the result stays in device
   * memory; each pull encodes the next batchSize rows as QMCB on the device
   * into pinned memory (an external ArrayBuffer, no copy) and yields them as rows.
   */
  async *cursor(desc, batchSize = 1000) {
    const left = this.tables.get(desc.table);
    if (!left) throw qmError('TABLE_NOT_FOUND', 'table ' + desc.table + ' is not registered');
    const right = desc.join ? this.tables.get(desc.join.right.table) : null;
    if (desc.join && !right) throw qmError('TABLE_NOT_FOUND', 'table ' + desc.join.right.table + ' is not registered');
    const plan = addon.planCompile(this.ctx, encodePlan(desc));
    let res = null;
    try {
      res = addon.exec(plan, left, right);
      const n = addon.resultRows(res);
      for (let r0 = 0; r0 < n; r0 += batchSize) {
        yield qmcbToRows(decodeQMCB(addon.resultQmcbRange(res, r0, batchSize)));
      }
    } finally {
      if (res) addon.resultClose(res);
      addon.planClose(plan);
    }
  }

  /** QueryExecutor.execute -> QueryResult-shaped object (src/types.ts:262-295). */
  execute(desc) {
    const t0 = Date.now();
    const r = this._run(desc);
    const batch = decodeQMCB(r.qmcb);
    const rows = qmcbToRows(batch);
    return { rows, rowCount: rows.length, columns: batch.columns.map((c) => c.name), columnarData: r.qmcb,
      bytesRead: r.stats.bytesAlgorithmic, pagesSkipped: r.stats.pagesSkipped, durationMs: Date.now() - t0, stats: r.stats };
  }

  /**
   * Many aggregate queries over one registered table in fused launches; the
   * Promise settles when the batch is done (the event loop runs meanwhile).
   */
  async executeBatch(descs) {
    if (!descs.length) return [];
    const name = descs[0].table;
    const table = this.tables.get(name);
    if (!table) throw qmError('TABLE_NOT_FOUND', 'table ' + name + ' is not registered');
    if (descs.some((d) => d.table !== name)) throw qmError('INVALID_ARGUMENT', 'a batch runs over one table');
    const t0 = Date.now();
    const plans = descs.map((d) => addon.planCompile(this.ctx, encodePlan(d)));
    try {
      const results = await addon.batchSubmit(this.ctx, plans, table);
      return results.map((res) => {
        try {
          const qmcb = addon.resultQmcb(res);
          const stats = addon.resultStats(res);
          const batch = decodeQMCB(qmcb);
          const rows = qmcbToRows(batch);
          return { rows, rowCount: rows.length, columns: batch.columns.map((c) => c.name), columnarData: qmcb,
            bytesRead: stats.bytesAlgorithmic, pagesSkipped: stats.pagesSkipped, durationMs: Date.now() - t0, stats };
        } finally { addon.resultClose(res); }
      });
    } finally { plans.forEach((p) => addon.planClose(p)); }
  }

  count(desc) { return this.execute(Object.assign({}, desc, { aggregates: [{ fn: 'count', column: '*' }], groupBy: [] })).rows[0]['count_*']; }

  /** A pull-based Operator over the device result: next() / nextColumnar() / close(). */
  toOperator(desc, batchSize) {
    const self = this;
    const size = batchSize || 4096;
    let rows = null;
    let pos = 0;
    let closed = false;
    const ensure = () => { if (rows === null) rows = self.execute(desc).rows; };
    return {
      next() {
        if (closed) return Promise.resolve(null);
        ensure();
        if (pos >= rows.length) return Promise.resolve(null);
        const out = rows.slice(pos, pos + size); pos += out.length;
        return Promise.resolve(out);
      },
      close() { closed = true; rows = []; return Promise.resolve(); },
    };
  }

  close() {
    for (const t of this.tables.values()) addon.tableDrop(t);
    this.tables.clear();
    addon.shutdown(this.ctx);
  }
}

/** drainPipeline (src/operators.ts:3155-3179): pull until null, close in finally. */
async function drainPipeline(op) {
  const out = [];
  try {
    for (;;) { const b = await op.next(); if (b === null) break; for (const r of b) out.push(r); }
  } finally { await op.close(); }
  return out;
}

/** The builder subset of the path (src/client.ts:176-312). */
class DataFrame {
  constructor(executor, desc) { this.executor = executor; this.desc = desc; }
  _with(patch) { return new DataFrame(this.executor, Object.assign({}, this.desc, patch)); }
  filter(column, op, value) { return this._with({ filters: (this.desc.filters || []).concat([{ column, op, value }]) }); }
  where(column, op, value) { return this.filter(column, op, value); }
  whereBetween(column, lo, hi) { return this.filter(column, 'between', [lo, hi]); }
  whereOr(...groups) { return this._with({ filterGroups: (this.desc.filterGroups || []).concat([groups.flat()]) }); }
  groupBy(...cols) { return this._with({ groupBy: (this.desc.groupBy || []).concat(cols) }); }
  aggregate(fn, column, alias) { return this._with({ aggregates: (this.desc.aggregates || []).concat([{ fn, column, alias }]) }); }
  select(...cols) { return this._with({ projections: cols }); }
  sort(column, direction) { return this._with({ sortColumn: column, sortDirection: direction || 'asc' }); }
  orderBy(column, direction) { return this.sort(column, direction); }
  limit(n) { return this._with({ limit: n }); }
  offset(n) { return this._with({ offset: n }); }
  join(other, on) { return this._with({ join: { right: other.desc, leftKey: on.left, rightKey: on.right, type: on.type || 'inner' } }); }
  computed(alias, a, b) { return this._with({ computedColumns: (this.desc.computedColumns || []).concat([[alias, a, b]]) }); }
  exec() { return this.executor.execute(this.desc); }
  collect() { return this.exec(); }
  toOperator(batchSize) { return this.executor.toOperator(this.desc, batchSize); }
}

class QueryMode {
  static gpu(device, hbmBudgetBytes) { return new QueryMode(new GpuExecutor(device, hbmBudgetBytes)); }
  constructor(executor) { this.executor = executor; }
  registerTable(name, columns, nulls) { return this.executor.registerTable(name, columns, nulls); }
  table(name) { return new DataFrame(this.executor, { table: name, filters: [], projections: [] }); }
  close() { this.executor.close(); }
}

module.exports = { addon, encodePlan, decodeQMCB, qmcbToRows, GpuExecutor, DataFrame, QueryMode, drainPipeline,
  DTYPE };
