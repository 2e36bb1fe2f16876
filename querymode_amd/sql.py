"""SQL text -> QueryDescriptor for the device path.

Covers the statement shapes the filter -> group-by -> join -> top-K path
executes, with the reference compiler's lowering rules
(src/sql/compiler.ts:39-140, 287-310, 365-531, 563-615):

  SELECT <*, columns, aggregates [AS alias]>
  FROM t [alias] [[INNER] JOIN r [alias] ON x.k = y.k | USING (k)]
  [WHERE <predicate>] [GROUP BY cols [HAVING <predicate>]]
  [ORDER BY col [ASC|DESC], ...] [LIMIT n] [OFFSET m]

* SELECT DISTINCT, HAVING and multi-column ORDER BY run as device stages
  after the aggregate / row selection, with the reference wrapper's
  semantics (src/sql/executor.ts:30-170): HAVING is evaluated with the SQL
  evaluator's comparisons and three-valued logic, and aggregates only HAVING
  names are computed but not returned (compiler.ts:159-164, executor.ts:143-160);
  a multi-column ORDER BY is a stable sort with NULLs last, then OFFSET /
  LIMIT; DISTINCT keeps the first row of each key (compiler.ts:154-157).

* Aggregates: COUNT(*), COUNT(col), SUM/AVG/MIN/MAX(col); an argument
  ``a * b`` is keyed ``multiply(a,b)`` (src/sql/ast.ts:129-149) and runs as a
  fused product aggregate.
* WHERE: a conjunction of simple predicates becomes ``filters``; a
  disjunction of conjunctions becomes ``filterGroups`` (compiler.ts:365-498).
  Where the reference falls back to a row-wise evaluator (a conjunction that
  contains an OR, NOT over AND/OR), this front end normalises instead: NOT is
  pushed to the leaves (De Morgan; NOT of a comparison flips the operator, so
  NULL rows still never match, as in a SQL WHERE) and the predicate is put in
  AND-of-(OR-of-AND) form, with at most one OR factor expanded into
  ``filterGroups``. A predicate that does not fit raises UNSUPPORTED — there
  is no row-wise CPU evaluator behind the device path.
* Literals: integers and decimals are JS numbers (compiler.ts:533-544);
  comparisons against column pairs, NULL literals, LIKE and strings are not on
  the numeric device path and raise UNSUPPORTED.
* With GROUP BY, the output columns are the group columns followed by the
  aggregate names (compiler.ts:80-88).
"""
from __future__ import annotations

import re
from dataclasses import dataclass

from .errors import QueryModeError
from .plan import AggregateOp, FilterOp, JoinDescriptor, QueryDescriptor

_KEYWORDS = {
    "SELECT", "FROM", "WHERE", "GROUP", "BY", "ORDER", "ASC", "DESC", "LIMIT", "OFFSET", "AND", "OR", "NOT",
    "BETWEEN", "IN", "IS", "NULL", "AS", "JOIN", "INNER", "LEFT", "RIGHT", "FULL", "OUTER", "ON", "USING",
    "LIKE", "TRUE", "FALSE", "DISTINCT", "CROSS", "HAVING",
}
_AGG_FNS = {"COUNT", "SUM", "AVG", "MIN", "MAX", "COUNT_DISTINCT", "STDDEV", "VARIANCE", "MEDIAN", "PERCENTILE"}
_TOKEN = re.compile(r"""
    (?P<ws>\s+)
  | (?P<num>\d+\.\d*(?:[eE][+-]?\d+)?|\.\d+(?:[eE][+-]?\d+)?|\d+(?:[eE][+-]?\d+)?)
  | (?P<str>'(?:[^']|'')*')
  | (?P<qid>"(?:[^"]|"")*")
  | (?P<id>[A-Za-z_][A-Za-z0-9_]*)
  | (?P<op><=|>=|<>|!=|=|<|>|\(|\)|,|\*|\.|-|\+|;)
""", re.VERBOSE)
MAX_SAFE = 2 ** 53 - 1
MAX_DEVICE_PREDS = 16


def _err(msg: str, code: str = "INVALID_FORMAT") -> QueryModeError:
    return QueryModeError(code, msg)


@dataclass
class _Tok:
    kind: str  # num, str, id, kw, op, end
    text: str
    pos: int


def _lex(sql: str) -> list[_Tok]:
    out, pos = [], 0
    while pos < len(sql):
        m = _TOKEN.match(sql, pos)
        if not m:
            raise _err(f"SQL: unexpected character {sql[pos]!r} at {pos}")
        kind = m.lastgroup
        text = m.group(kind)
        if kind == "id" and text.upper() in _KEYWORDS:
            out.append(_Tok("kw", text.upper(), pos))
        elif kind == "qid":
            out.append(_Tok("id", text[1:-1].replace('""', '"'), pos))
        elif kind != "ws":
            out.append(_Tok(kind, text, pos))
        pos = m.end()
    out.append(_Tok("end", "", len(sql)))
    return out


# ---- expression AST (tuples) ----
# ("col", name) ("lit", value) ("null",) ("str", s) ("bool", b)
# ("cmp", op, left, right) ("and", a, b) ("or", a, b) ("not", a)
# ("between", expr, lo, hi, negated) ("in", expr, [values], negated)
# ("isnull", expr, negated) ("like", expr, pattern, negated) ("mul", a, b) ("call", NAME, [args], distinct)
# ("star",)


class _Parser:
    def __init__(self, sql: str):
        self.t = _lex(sql)
        self.i = 0

    def peek(self, k: int = 0) -> _Tok:
        return self.t[min(self.i + k, len(self.t) - 1)]

    def take(self) -> _Tok:
        tok = self.t[self.i]
        self.i += 1
        return tok

    def accept(self, kind: str, text: str | None = None) -> _Tok | None:
        tok = self.peek()
        if tok.kind == kind and (text is None or tok.text == text):
            return self.take()
        return None

    def expect(self, kind: str, text: str | None = None) -> _Tok:
        tok = self.accept(kind, text)
        if tok is None:
            got = self.peek()
            raise _err(f"SQL: expected {text or kind} at {got.pos}, got {got.text!r}")
        return tok

    def ident(self) -> str:
        tok = self.peek()
        if tok.kind != "id":
            raise _err(f"SQL: expected an identifier at {tok.pos}, got {tok.text!r}")
        return self.take().text

    # column reference: name or qualifier.name
    def column(self) -> tuple:
        name = self.ident()
        if self.accept("op", "."):
            return ("col", self.ident(), name)
        return ("col", name, None)

    # ---- statement ----
    def select(self) -> dict:
        self.expect("kw", "SELECT")
        distinct = bool(self.accept("kw", "DISTINCT"))
        items = [self.select_item()]
        while self.accept("op", ","):
            items.append(self.select_item())
        self.expect("kw", "FROM")
        frm = self.table_ref()
        join = None
        if self.peek().kind == "kw" and self.peek().text in ("JOIN", "INNER", "LEFT", "RIGHT", "FULL", "CROSS"):
            join = self.join_clause()
        where = None
        if self.accept("kw", "WHERE"):
            where = self.expr()
        group = []
        if self.accept("kw", "GROUP"):
            self.expect("kw", "BY")
            group.append(self.column()[1])
            while self.accept("op", ","):
                group.append(self.column()[1])
        having = None
        if self.accept("kw", "HAVING"):
            having = self.expr()
        order = []
        if self.accept("kw", "ORDER"):
            self.expect("kw", "BY")
            order.append(self.order_item())
            while self.accept("op", ","):
                order.append(self.order_item())
        limit = offset = None
        if self.accept("kw", "LIMIT"):
            limit = int(self.expect("num").text)
        if self.accept("kw", "OFFSET"):
            offset = int(self.expect("num").text)
        self.accept("op", ";")
        if self.peek().kind != "end":
            raise _err(f"SQL: unexpected {self.peek().text!r} at {self.peek().pos}")
        return {"items": items, "from": frm, "join": join, "where": where, "group": group, "order": order,
                "limit": limit, "offset": offset, "distinct": distinct, "having": having}

    def select_item(self):
        if self.accept("op", "*"):
            return (("star",), None)
        e = self.value_expr()
        alias = None
        if self.accept("kw", "AS"):
            alias = self.ident()
        elif self.peek().kind == "id":
            alias = self.ident()
        return (e, alias)

    def table_ref(self) -> tuple[str, str | None]:
        name = self.ident()
        alias = None
        if self.accept("kw", "AS"):
            alias = self.ident()
        elif self.peek().kind == "id":
            alias = self.ident()
        return name, alias

    def join_clause(self) -> dict:
        jtype = "inner"
        tok = self.peek()
        if tok.text in ("LEFT", "RIGHT", "FULL", "CROSS"):
            jtype = self.take().text.lower()
            self.accept("kw", "OUTER")
        else:
            self.accept("kw", "INNER")
        self.expect("kw", "JOIN")
        right = self.table_ref()
        on = using = None
        if self.accept("kw", "ON"):
            a = self.column()
            self.expect("op", "=")
            b = self.column()
            on = (a, b)
        elif self.accept("kw", "USING"):
            self.expect("op", "(")
            using = self.ident()
            self.expect("op", ")")
        return {"type": jtype, "right": right, "on": on, "using": using}

    def order_item(self):
        col = self.column()[1]
        direction = "asc"
        if self.accept("kw", "DESC"):
            direction = "desc"
        else:
            self.accept("kw", "ASC")
        return col, direction

    # ---- predicates ----
    def expr(self):
        e = self.and_expr()
        while self.accept("kw", "OR"):
            e = ("or", e, self.and_expr())
        return e

    def and_expr(self):
        e = self.not_expr()
        while self.accept("kw", "AND"):
            e = ("and", e, self.not_expr())
        return e

    def not_expr(self):
        if self.accept("kw", "NOT"):
            return ("not", self.not_expr())
        return self.predicate()

    def predicate(self):
        if self.peek().kind == "op" and self.peek().text == "(":
            # parenthesised predicate or value expression
            save = self.i
            self.take()
            try:
                e = self.expr()
                self.expect("op", ")")
                if not (self.peek().kind == "op" and self.peek().text in ("=", "!=", "<>", "<", "<=", ">", ">=")):
                    return e
            except QueryModeError:
                pass
            self.i = save
        left = self.value_expr()
        tok = self.peek()
        if tok.kind == "op" and tok.text in ("=", "!=", "<>", "<", "<=", ">", ">="):
            self.take()
            op = {"=": "eq", "!=": "ne", "<>": "ne", "<": "lt", "<=": "le", ">": "gt", ">=": "ge"}[tok.text]
            return ("cmp", op, left, self.value_expr())
        negated = bool(self.accept("kw", "NOT"))
        if self.accept("kw", "BETWEEN"):
            lo = self.value_expr()
            self.expect("kw", "AND")
            hi = self.value_expr()
            return ("between", left, lo, hi, negated)
        if self.accept("kw", "IN"):
            self.expect("op", "(")
            vals = [self.value_expr()]
            while self.accept("op", ","):
                vals.append(self.value_expr())
            self.expect("op", ")")
            return ("in", left, vals, negated)
        if self.accept("kw", "LIKE"):
            return ("like", left, self.value_expr(), negated)
        if negated:
            raise _err(f"SQL: expected BETWEEN, IN or LIKE after NOT at {self.peek().pos}")
        if self.accept("kw", "IS"):
            neg = bool(self.accept("kw", "NOT"))
            self.expect("kw", "NULL")
            return ("isnull", left, neg)
        if left[0] == "bool":
            return left
        raise _err(f"SQL: expected a comparison at {tok.pos}, got {tok.text!r}")

    # ---- value expressions: products of primaries ----
    def value_expr(self):
        e = self.primary()
        while self.peek().kind == "op" and self.peek().text == "*":
            self.take()
            e = ("mul", e, self.primary())
        return e

    def primary(self):
        tok = self.peek()
        if tok.kind == "op" and tok.text == "-":
            self.take()
            v = self.primary()
            if v[0] != "lit":
                raise _err(f"SQL: unary minus needs a number at {tok.pos}")
            return ("lit", -v[1])
        if tok.kind == "op" and tok.text == "+":
            self.take()
            return self.primary()
        if tok.kind == "num":
            self.take()
            txt = tok.text
            if re.fullmatch(r"\d+", txt):
                v = int(txt)
                return ("lit", v if v <= MAX_SAFE else float(v))  # JS number literal
            return ("lit", float(txt))
        if tok.kind == "str":
            self.take()
            return ("str", tok.text[1:-1].replace("''", "'"))
        if tok.kind == "kw" and tok.text == "NULL":
            self.take()
            return ("null",)
        if tok.kind == "kw" and tok.text in ("TRUE", "FALSE"):
            self.take()
            return ("bool", tok.text == "TRUE")
        if tok.kind == "op" and tok.text == "(":
            self.take()
            e = self.value_expr()
            self.expect("op", ")")
            return e
        if tok.kind == "id":
            if self.peek(1).kind == "op" and self.peek(1).text == "(":
                name = self.take().text.upper()
                self.take()
                distinct = bool(self.accept("kw", "DISTINCT"))
                args = []
                if self.accept("op", "*"):
                    args.append(("star",))
                elif not (self.peek().kind == "op" and self.peek().text == ")"):
                    args.append(self.value_expr())
                    while self.accept("op", ","):
                        args.append(self.value_expr())
                self.expect("op", ")")
                return ("call", name, args, distinct)
            return self.column()
        raise _err(f"SQL: unexpected {tok.text!r} at {tok.pos}")


# ---- lowering ----

def _arg_key(e) -> str:
    """Aggregate argument key (src/sql/ast.ts:129-149)."""
    k = e[0]
    if k == "star":
        return "*"
    if k == "col":
        return e[1]
    if k == "lit":
        return repr(e[1]) if isinstance(e[1], float) and not e[1].is_integer() else str(int(e[1]) if isinstance(e[1], float) else e[1])
    if k == "mul":
        return f"multiply({_arg_key(e[1])},{_arg_key(e[2])})"
    raise _err("SQL: unsupported aggregate argument", "UNSUPPORTED")


_FLIP = {"lt": "gt", "le": "ge", "gt": "lt", "ge": "le", "eq": "eq", "ne": "ne"}
_OPNAME = {"eq": "eq", "ne": "neq", "lt": "lt", "le": "lte", "gt": "gt", "ge": "gte"}


def _lit_value(e):
    if e[0] == "lit":
        return e[1]
    if e[0] == "null":
        raise _err("SQL: NULL literals in predicates are not on the device path (use IS NULL)", "UNSUPPORTED")
    if e[0] in ("str", "bool"):
        raise _err("SQL: string/boolean literals are not on the numeric device path", "UNSUPPORTED")
    raise _err("SQL: predicate operand must be a literal", "UNSUPPORTED")


def _leaf(e, neg: bool = False) -> FilterOp:
    """One simple predicate (compiler.ts:372-451, 501-531); `neg` is set only
    for NOT BETWEEN, the one NOT the filter path takes."""
    k = e[0]
    if k == "cmp":
        op, a, b = e[1], e[2], e[3]
        if a[0] == "col" and b[0] != "col":
            col, val = a[1], _lit_value(b)
        elif b[0] == "col" and a[0] != "col":
            col, val, op = b[1], _lit_value(a), _FLIP[op]
        else:
            raise _err("SQL: column-to-column comparisons are not on the device path", "UNSUPPORTED")
        return FilterOp(col, _OPNAME[op], val)
    if k == "between":
        if e[1][0] != "col":
            raise _err("SQL: BETWEEN needs a column", "UNSUPPORTED")
        negated = e[4] != neg
        return FilterOp(e[1][1], "not_between" if negated else "between", [_lit_value(e[2]), _lit_value(e[3])])
    if k == "in":
        if e[1][0] != "col":
            raise _err("SQL: IN needs a column", "UNSUPPORTED")
        return FilterOp(e[1][1], "not_in" if e[3] else "in", [_lit_value(v) for v in e[2]])
    if k == "isnull":
        if e[1][0] != "col":
            raise _err("SQL: IS NULL needs a column", "UNSUPPORTED")
        return FilterOp(e[1][1], "is_not_null" if e[2] else "is_null", 0)
    if k == "like":
        raise _err("SQL: LIKE needs string columns, which are not on the device path", "UNSUPPORTED")
    raise _err(f"SQL: unsupported predicate {k}", "UNSUPPORTED")


def _flatten(e, out: list[FilterOp]) -> bool:
    """tryFlattenFilters (compiler.ts:365-466): an AND tree of simple
    predicates, where the only NOT taken is NOT BETWEEN (NOT LIKE is a string
    predicate). False for anything else: OR, NOT over a comparison, IN or IS
    NULL, NOT over AND/OR, a doubled NOT."""
    k = e[0]
    if k == "and":
        left = _flatten(e[1], out)
        right = _flatten(e[2], out)
        return left and right
    if k in ("cmp", "between", "in", "isnull", "like"):
        out.append(_leaf(e))
        return True
    if k == "not" and e[1][0] == "between" and not e[1][4]:
        out.append(_leaf(e[1], True))
        return True
    if k == "not" and e[1][0] == "like":
        out.append(_leaf(e[1]))
    return False


def _or_branches(e, out: list) -> list:
    """collectOrBranches (compiler.ts:487-495)."""
    if e[0] == "or":
        _or_branches(e[1], out)
        _or_branches(e[2], out)
    else:
        out.append(e)
    return out


_ROW_BY_ROW = ("SQL: this WHERE clause is evaluated row by row by the SQL evaluator, not on the filter path "
               "(NOT over a comparison, IN or IS NULL, NOT over AND/OR, or an AND that holds an OR); "
               "it is not on the device path")


def lower_where(where) -> tuple[list[FilterOp], list[list[FilterOp]]]:
    """WHERE -> (filters, filterGroups) exactly where the reference's SQL
    compiler pushes the clause to the filter path (compiler.ts:90-109): an AND
    tree of simple predicates becomes `filters`; failing that, a top-level OR
    whose every branch is such a tree becomes `filterGroups`. Every other
    clause goes to the reference's row-by-row evaluator (executor.ts:89-91),
    whose comparisons differ from the filter path's (NaN sorts greatest,
    integer columns compare exactly against fractional literals,
    evaluator.ts:169-200), so it raises UNSUPPORTED here rather than run with
    the filter path's meaning."""
    filters: list[FilterOp] = []
    groups: list[list[FilterOp]] = []
    if not _flatten(where, filters):
        filters = []
        branches = _or_branches(where, [])
        if len(branches) < 2:
            raise _err(_ROW_BY_ROW, "UNSUPPORTED")
        for br in branches:
            g: list[FilterOp] = []
            if not _flatten(br, g) or not g:
                raise _err(_ROW_BY_ROW, "UNSUPPORTED")
            groups.append(g)
    if len(filters) + sum(len(g) for g in groups) > MAX_DEVICE_PREDS:
        raise _err("SQL: WHERE clause has more predicates than the device path takes", "UNSUPPORTED")
    return filters, groups


def _agg_of(e) -> tuple[str, str]:
    """(fn, column key) of an aggregate call (compiler.ts:287-310)."""
    fn = e[1].lower()
    if fn == "count" and e[3]:
        fn = "count_distinct"
    col = "*" if not e[2] or e[2][0][0] == "star" else _arg_key(e[2][0])
    return fn, col


def _having_operand(e, aggs: list[AggregateOp]):
    """extractAggregatesFromExpr + rewriteAggregatesAsColumns
    (compiler.ts:254-285, evaluator.ts:263-275): an aggregate call becomes the
    column of its output (the SELECT alias when one matches), and an
    aggregate that only HAVING uses is added to the query's aggregates."""
    if e[0] == "call" and e[1] in _AGG_FNS:
        fn, col = _agg_of(e)
        match = next((a for a in aggs if a.fn == fn and a.column == col), None)
        if match is None:
            match = AggregateOp(fn, col)
            aggs.append(match)
        return ("col", match.out_name())
    return e


def lower_having(e, aggs: list[AggregateOp]) -> tuple:
    """HAVING -> the device's HAVING tree (plan.HAVING_NODES), evaluated with the
    SQL evaluator's semantics (src/sql/executor.ts:108-112)."""
    k = e[0]
    if k in ("and", "or"):
        return (k, lower_having(e[1], aggs), lower_having(e[2], aggs))
    if k == "not":
        return ("not", lower_having(e[1], aggs))
    if k == "cmp":
        op, a, b = e[1], _having_operand(e[2], aggs), _having_operand(e[3], aggs)
        if a[0] == "col" and b[0] == "col":
            return ("cmp_col", op, a[1], b[1])
        if a[0] == "col":
            return ("cmp", op, a[1], _lit_value(b))
        if b[0] == "col":
            return ("cmp", _FLIP[op], b[1], _lit_value(a))
        raise _err("SQL: HAVING comparison needs a column or aggregate", "UNSUPPORTED")
    if k == "between":
        a = _having_operand(e[1], aggs)
        if a[0] != "col":
            raise _err("SQL: HAVING BETWEEN needs a column or aggregate", "UNSUPPORTED")
        return ("between", a[1], _lit_value(e[2]), _lit_value(e[3]), bool(e[4]))
    if k == "in":
        a = _having_operand(e[1], aggs)
        if a[0] != "col":
            raise _err("SQL: HAVING IN needs a column or aggregate", "UNSUPPORTED")
        return ("in", a[1], [None if v[0] == "null" else _lit_value(v) for v in e[2]], bool(e[3]))
    if k == "isnull":
        a = _having_operand(e[1], aggs)
        if a[0] != "col":
            raise _err("SQL: HAVING IS NULL needs a column or aggregate", "UNSUPPORTED")
        return ("isnull", a[1], bool(e[2]))
    raise _err(f"SQL: unsupported HAVING term {k}", "UNSUPPORTED")


def sql_to_descriptor(sql: str) -> QueryDescriptor:
    st = _Parser(sql).select()
    table, talias = st["from"]
    desc = QueryDescriptor(table=table)
    aggs: list[AggregateOp] = []
    proj: list[str] = []
    for e, alias in st["items"]:
        if e[0] == "star":
            continue
        if e[0] == "call" and e[1] in _AGG_FNS:
            fn = e[1].lower()
            if fn == "count" and e[3]:
                fn = "count_distinct"
            col = "*" if not e[2] or e[2][0][0] == "star" else _arg_key(e[2][0])
            aggs.append(AggregateOp(fn, col, alias))
            continue
        if e[0] == "col":
            proj.append(e[1])
            continue
        raise _err("SQL: computed select expressions are not on the device path", "UNSUPPORTED")
    desc.aggregates = aggs
    desc.groupBy = list(st["group"])
    if aggs and st["group"]:
        proj = list(st["group"]) + [a.out_name() for a in aggs]
    elif aggs:
        proj = []
    desc.projections = proj
    if st["join"] is not None:
        j = st["join"]
        rname, ralias = j["right"]
        if j["type"] != "inner":
            raise _err(f"SQL: {j['type'].upper()} JOIN is not on the device path (inner joins only)", "UNSUPPORTED")
        if j["on"] is not None:
            a, b = j["on"]
            rq = {rname, ralias} - {None}
            if a[2] is not None and a[2] in rq:
                lk, rk = b[1], a[1]
            else:
                lk, rk = a[1], b[1]
        elif j["using"] is not None:
            lk = rk = j["using"]
        else:
            raise _err("SQL: JOIN needs ON a = b or USING (col)")
        desc.join = JoinDescriptor(QueryDescriptor(table=rname), lk, rk, "inner")
    if st["where"] is not None:
        desc.filters, desc.filterGroups = lower_where(st["where"])
    if st["having"] is not None:
        if not st["group"]:
            raise _err("SQL: HAVING needs GROUP BY", "UNSUPPORTED")
        desc.having = lower_having(st["having"], desc.aggregates)
    if st["distinct"]:
        desc.distinct = list(proj)  # compiler.ts:154-157
    if st["order"]:
        if len(st["order"]) > 1:
            desc.orderBy = list(st["order"])  # the wrapper's multi-column sort (compiler.ts:124-131)
        else:
            desc.sortColumn, desc.sortDirection = st["order"][0]
    desc.limit = st["limit"]
    desc.offset = st["offset"]
    return desc


__all__ = ["sql_to_descriptor", "lower_where"]
