#!/usr/bin/env python3
"""Reverse- and forward-mode source transformation of the single-assignment
flux listings (*.ssa) into C++ (mhd2d_face_ad.h).

Each listing statement `x = expr` (Python expression syntax, which parses
+ - * / and unary minus with C's precedence and left associativity) is
lowered to three-address operations in the expression's own evaluation
order, so the emitted primal rounds exactly as the written C expression
does.  For every listing three functions are emitted:

  NAME(in, par, out)                  primal
  NAME_tlm(in, in_d, par, out, out_d) tangent-linear (one direction)
  NAME_adj(in, par, out_b, in_b)      adjoint: in_b = (d out / d in)^T out_b

  NAME_rep(in, in_p, par, out, out_p) decision replay: out_p is the primal
                                      at in_p with every branch decision
                                      taken at in (see replay_lines)

Elementary operations and their derivatives:
  a+b, a-b, a*b, a/b, -a       the usual rules
  sqrt(a)                      1/(2 sqrt a), taken as 0 where sqrt a = 0
  fabs(a)                      sign(a) (0 at 0)
  tmin(a, b) = b < a ? b : a   routes to the selected argument
  tmax(a, b) = a < b ? b : a   routes to the selected argument
  sgn(a), ind(condition)       piecewise constant: derivative 0
  minmod(a, b)                 expanded as s = sgn(a); s * tmax(tmin(a s, s b), eps)
                               (a s = |a|)
The branch decisions (minmod, tmin/tmax, the HLLD region indicators) are
taken from the primal values, which the adjoint recomputes first, so the
adjoint differentiates exactly the branch the forward step took.

Usage: gen_face_ad.py OUT.h LISTING.ssa[:NAME] ...
"""
import ast
import re
import os
import sys


class Lowering:
    def __init__(self, params):
        self.params = set(params)
        self.ops = []          # (dst, op, args) ; args are names or float literals
        self.ntmp = 0

    def tmp(self):
        self.ntmp += 1
        return f"t{self.ntmp}"

    def operand(self, node):
        if isinstance(node, ast.Constant):
            return repr(float(node.value))
        if isinstance(node, ast.Name):
            return node.id
        return self.expr(node)

    def emit(self, op, args, dst=None):
        dst = dst or self.tmp()
        self.ops.append((dst, op, args))
        return dst

    def expr(self, node, dst=None):
        if isinstance(node, ast.BinOp):
            a = self.operand(node.left)
            b = self.operand(node.right)
            op = {ast.Add: "add", ast.Sub: "sub", ast.Mult: "mul", ast.Div: "div"}[type(node.op)]
            return self.emit(op, [a, b], dst)
        if isinstance(node, ast.UnaryOp) and isinstance(node.op, ast.USub):
            return self.emit("neg", [self.operand(node.operand)], dst)
        if isinstance(node, ast.Call):
            fn = node.func.id
            if fn == "ind":
                return self.emit("ind", [self.cond(node.args[0])], dst)
            if fn == "minmod":
                a = self.operand(node.args[0])
                b = self.operand(node.args[1])
                s = self.emit("sgn", [a])
                # |a| as a * sgn(a): the same value (bar the sign of a zero,
                # which the limiter's eps floor absorbs, so the minmod is
                # bitwise the reference's), and its derivative reads s, which
                # the reverse pass holds anyway, instead of testing a again
                ax = self.emit("mul", [a, s])
                sy = self.emit("mul", [s, b])
                m = self.emit("tmin", [ax, sy])
                mm = self.emit("tmax", [m, "eps"])
                return self.emit("mul", [s, mm], dst)
            args = [self.operand(x) for x in node.args]
            return self.emit(fn, args, dst)
        if isinstance(node, (ast.Name, ast.Constant)):
            return self.emit("copy", [self.operand(node)], dst)
        raise ValueError(f"unsupported expression {ast.dump(node)}")

    def cond(self, node):
        if isinstance(node, ast.BoolOp):
            j = " && " if isinstance(node.op, ast.And) else " || "
            return "(" + j.join(self.cond(v) for v in node.values) + ")"
        if isinstance(node, ast.Compare) and len(node.ops) == 1:
            sym = {ast.Lt: "<", ast.LtE: "<=", ast.Gt: ">", ast.GtE: ">="}[type(node.ops[0])]
            return f"({self.operand(node.left)} {sym} {self.operand(node.comparators[0])})"
        raise ValueError(f"unsupported condition {ast.dump(node)}")


def parse(path):
    """Returns params, inputs, outputs and ops; parse.guards holds the guard
    (`guard:` lines) each op is under."""
    params, inputs, outputs, stmts = [], [], [], []
    guard = None
    for line in open(path):
        line = line.split("#", 1)[0].strip()
        if not line:
            continue
        if line.startswith("params:"):
            params = line.split(":", 1)[1].split()
        elif line.startswith("inputs:"):
            inputs = line.split(":", 1)[1].split()
        elif line.startswith("outputs:"):
            outputs = line.split(":", 1)[1].split()
        elif line.startswith("guard:"):
            g = line.split(":", 1)[1].strip()
            guard = None if g == "none" else g
        else:
            stmts.append((guard, ast.parse(line).body[0]))
    low = Lowering(params)
    guards = []
    for gd, st in stmts:
        n0 = len(low.ops)
        low.expr(st.value, dst=st.targets[0].id)
        guards += [gd] * (len(low.ops) - n0)
    parse.guards = guards
    return params, inputs, outputs, low.ops


C_OP = {"dlim": "{0}", "add": "{0} + {1}", "sub": "{0} - {1}", "mul": "{0} * {1}", "div": "{0} / {1}", "neg": "-{0}",
        "copy": "{0}", "sqrt": "sqrt({0})", "fabs": "fabs({0})", "tmin": "(({1}) < ({0}) ? ({1}) : ({0}))",
        "tmax": "(({0}) < ({1}) ? ({1}) : ({0}))", "sgn": "(double)(({0} > 0.0) - ({0} < 0.0))",
        "ind": "(double){0}"}


def is_lit(a):
    return a[0].isdigit() or a[0] == "-"


def groups(guards):
    """Runs of consecutive ops with the same guard: (guard, [op indices])."""
    out = []
    for n, g in enumerate(guards):
        if out and out[-1][0] == g:
            out[-1][1].append(n)
        else:
            out.append((g, [n]))
    return out


# The adjoint's own primal pass: values that no branch depends on only set
# the point the derivatives are taken at, so there division and square root
# use the refined hardware estimates.  Everything a comparison or a guard
# reads, transitively, keeps IEEE arithmetic, so every branch replays the
# forward step's decision exactly.
C_OP_FAST = dict(C_OP, div="{0} * mhd_ad_rcp({1})", sqrt="mhd_ad_sqrt({0})")
BRANCH_OPS = ("tmin", "tmax", "sgn", "ind", "dlim")


def names_in(args):
    """Variable names an op reads: its operands, and the names inside the
    condition text of an ind op."""
    return {n for a in args if not is_lit(a) for n in re.findall(r"[A-Za-z_][A-Za-z_0-9]*", a)}


def branch_inputs(ops, guards):
    """Names some branch (comparison op or guard) depends on, transitively.
    dlim(x, limit, c) branches on c only."""
    need = set()
    for dst, op, args in ops:
        if op in BRANCH_OPS:
            need.update(names_in(args[2:] if op == "dlim" else args))
    for g in guards or []:
        if g:
            need.update(names_in([g]))
    for dst, op, args in reversed(ops):
        if dst in need:
            need.update(names_in(args))
    return need


def decision_bits(ops):
    """Bit (word, bit) of each tmin / tmax decision in the adjoint's packed
    decision words."""
    out = {}
    for dst, op, _ in ops:
        if op in ("tmin", "tmax"):
            k = len(out)
            out[dst] = (k // 32, k % 32)
    return out


def select_cond(op, a, b):
    """The condition under which tmin / tmax (a, b) selects b."""
    return f"({b}) < ({a})" if op == "tmin" else f"({a}) < ({b})"


def primal_lines(params, inputs, ops, guards=None, fast=False, decisions=None):
    """Primal statements.  With guards (tangent and adjoint only), an op
    under a guard is evaluated only where the guard holds (its result feeds
    nothing else there); the plain primal evaluates every op.  fast: the
    adjoint's primal arithmetic (C_OP_FAST off every branch's inputs).
    decisions: the adjoint's tmin / tmax decisions, recorded as bits of
    integer words (one VGPR holds 32 of them) for the reverse pass to test,
    instead of lane masks that would stay live across the face in scalar
    registers."""
    exact = branch_inputs(ops, guards) if fast else None

    def C(dst, op):
        return C_OP_FAST[op] if fast and dst not in exact else C_OP[op]

    def record(dst, op, args, indent):
        if decisions and dst in decisions:
            w, bit = decisions[dst]
            return [f"{indent}dec{w} |= ({select_cond(op, *args)}) ? {1 << bit}u : 0u;"]
        return []
    out = [f"    const double {p} = par[{i}];" for i, p in enumerate(params)]
    out += [f"    const double {v} = in[{i}];" for i, v in enumerate(inputs)]
    if decisions:
        out += [f"    unsigned dec{w} = 0u;" for w in sorted({w for w, _ in decisions.values()})]
    if guards is None:
        guards = [None] * len(ops)
    for g, idx in groups(guards):
        if g is None:
            for n in idx:
                dst, op, args = ops[n]
                out.append(f"    const double {dst} = {C(dst, op).format(*args)};")
                out += record(dst, op, args, "    ")
        else:
            out += [f"    double {ops[n][0]} = 0.0;" for n in idx]
            out.append(f"    if ({g}) {{")
            for n in idx:
                dst, op, args = ops[n]
                out.append(f"    {dst} = {C(dst, op).format(*args)};")
                out += record(dst, op, args, "    ")
            out.append("    }")
    return out


def tangent_lines(params, inputs, ops, guards=None):
    """Tangent statements.  Names whose tangent is identically zero (the
    parameters, sgn and ind, and ops of such names only) enter the rules as
    the literal 0.0, and terms multiplying it are left out: IEEE arithmetic
    cannot drop 0.0 * x or 0.0 + x by itself."""
    zero = set(params)
    d = lambda a: "0.0" if (is_lit(a) or a in zero) else f"{a}_d"
    out = []
    for i, v in enumerate(inputs):
        out.append(f"    const double {v}_d = in_d[{i}];")
    if guards is None:
        guards = [None] * len(ops)
    for g, idx in groups(guards):
        body = tangent_ops([ops[n] for n in idx], d, zero)
        if g is None:
            out += [f"    const double {x}" for x in body]
        else:
            out += [f"    double {ops[n][0]}_d = 0.0;" for n in idx]
            out.append(f"    if ({g}) {{")
            out += [f"    {x}" for x in body]
            out.append("    }")
    return out


def tangent_ops(ops, d, zero):
    """Forward-mode rules; a dst whose tangent comes out as 0.0 joins zero."""
    Z = "0.0"
    out = []
    for dst, op, args in ops:
        a = args[0] if args else None
        b = args[1] if len(args) > 1 else None
        da = d(a) if a is not None else Z
        db = d(b) if b is not None else Z
        if op == "add":
            terms = [x for x in (da, db) if x != Z]
            e = " + ".join(terms) if terms else Z
        elif op == "sub":
            e = da if db == Z else (f"-{db}" if da == Z else f"{da} - {db}")
        elif op == "mul":
            terms = ([f"{da} * {b}"] if da != Z else []) + ([f"{a} * {db}"] if db != Z else [])
            e = " + ".join(terms) if terms else Z
        elif op == "div":
            if db == Z:
                e = Z if da == Z else f"{da} * mhd_ad_rcp({b})"
            else:
                num = f"-{dst} * {db}" if da == Z else f"{da} - {dst} * {db}"
                e = f"({num}) * mhd_ad_rcp({b})"
        elif op == "neg":
            e = Z if da == Z else f"-{da}"
        elif op == "copy":
            e = da
        elif op == "sqrt":
            e = Z if da == Z else f"({dst} > 0.0 ? 0.5 * {da} * mhd_ad_rcp({dst}) : 0.0)"
        elif op == "fabs":
            e = Z if da == Z else f"({a} > 0.0 ? {da} : ({a} < 0.0 ? -{da} : 0.0))"
        elif op == "tmin":
            e = Z if da == Z and db == Z else f"(({b}) < ({a}) ? {db} : {da})"
        elif op == "tmax":
            e = Z if da == Z and db == Z else f"(({a}) < ({b}) ? {db} : {da})"
        elif op in ("sgn", "ind"):
            e = Z
        elif op == "dlim":
            e = Z if da == Z and db == Z else f"({args[2]} != 0.0 ? {db} : {da})"
        else:
            raise ValueError(op)
        if e == Z:
            zero.add(dst)
        out.append(f"{dst}_d = {e};")
    return out


def replay_lines(params, inputs, ops):
    """Primal of the perturbed inputs in_p with the branch decisions of the
    base inputs in (whose primal is evaluated first): tmin / tmax select the
    argument the base selected, fabs keeps the base sign, sgn and ind keep
    their base values, and dlim(x, limit, c) adds the limit's change to x's
    base value where c holds.  The map in_p -> out_p is smooth, equals the
    primal at in_p = in (but for the sign of a zero) and its derivative at
    in is the tangent-linear model's, so central differences of it check the
    tangent and the adjoint without the jumps the branches put into the
    primal itself."""
    cst = set(params)
    P = lambda a: a if (is_lit(a) or a in cst) else f"{a}_p"
    out = [f"    const double {v}_p = in_p[{i}];" for i, v in enumerate(inputs)]
    for dst, op, args in ops:
        a = args[0] if args else None
        b = args[1] if len(args) > 1 else None
        if op in ("add", "sub", "mul", "div", "neg", "copy", "sqrt"):
            e = C_OP[op].format(*[P(x) for x in args])
        elif op == "fabs":
            e = f"({a} > 0.0 ? {P(a)} : ({a} < 0.0 ? -{P(a)} : 0.0))"
        elif op == "tmin":
            e = f"(({b}) < ({a}) ? {P(b)} : {P(a)})"
        elif op == "tmax":
            e = f"(({a}) < ({b}) ? {P(b)} : {P(a)})"
        elif op in ("sgn", "ind"):
            e = dst
        elif op == "dlim":
            e = f"({args[2]} != 0.0 ? {a} + ({P(b)} - {b}) : {P(a)})"
        else:
            raise ValueError(op)
        out.append(f"    const double {dst}_p = {e};")
    return out


def reverse_op(dst, op, args, active, ren, first, decisions=None):
    """Adjoint statements of one three-address op; ren maps a primal name to
    the local (recomputed) copy its value is read from.  first holds the
    adjoints already written in program order: the first contribution to
    an adjoint is assigned rather than added to its 0.0, because IEEE
    arithmetic cannot fold 0.0 + x (it is +0.0 for x = -0.0)."""
    def acc(v, e):
        if v in first:
            return f"{v}_b += {e};"
        first.add(v)
        return f"{v}_b = {e};"
    out = []
    a = args[0] if args else None
    b = args[1] if len(args) > 1 else None
    t = f"{dst}_b"
    V = lambda x: ren.get(x, x)
    terms = []
    if op == "add":
        terms = [(a, t), (b, t)]
    elif op == "sub":
        terms = [(a, t), (b, f"-{t}")]
    elif op == "mul":
        terms = [(a, f"{t} * {V(b)}"), (b, f"{t} * {V(a)}")]
    elif op == "div":
        out.append(f"    const double {dst}_r = mhd_ad_rcp({V(b)});")
        terms = [(a, f"{t} * {dst}_r"), (b, f"-{t} * {V(dst)} * {dst}_r")]
    elif op == "neg":
        terms = [(a, f"-{t}")]
    elif op == "copy":
        terms = [(a, t)]
    elif op == "sqrt":
        terms = [(a, f"({V(dst)} > 0.0 ? 0.5 * {t} * mhd_ad_rcp({V(dst)}) : 0.0)")]
    elif op == "fabs":
        terms = [(a, f"({V(a)} > 0.0 ? {t} : ({V(a)} < 0.0 ? -{t} : 0.0))")]
    elif op in ("tmin", "tmax"):
        if decisions and dst in decisions:
            w, bit = decisions[dst]
            sel = f"(dec{w} & {1 << bit}u) != 0u"
        else:
            sel = select_cond(op, V(a), V(b))
        if active(a) and active(b):
            sb, sa = acc(b, t), acc(a, t)
            out.append(f"    if ({sel}) {sb} else {sa}")
        elif active(a):
            out.append(f"    if (!({sel})) {acc(a, t)}")
        elif active(b):
            out.append(f"    if ({sel}) {acc(b, t)}")
        return out
    elif op in ("sgn", "ind"):
        return out
    elif op == "dlim":
        c = V(args[2])
        if active(b):
            out.append(f"    if ({c} != 0.0) {acc(b, t)}")
        if active(a):
            out.append(f"    if ({c} == 0.0) {acc(a, t)}")
        return out
    else:
        raise ValueError(op)
    for v, e in terms:
        if active(v):
            out.append(f"    {acc(v, e)}")
    return out


def grad_live(ops, outputs):
    """Names whose adjoint can be nonzero: those that reach an output
    through differentiable operands (not through sgn, ind or the condition
    of dlim).  Reverse statements into any other name would add terms that
    are zero in value but that IEEE arithmetic cannot fold away."""
    live = set(outputs)
    for dst, op, args in reversed(ops):
        if dst in live:
            if op in ("sgn", "ind"):
                continue
            live.update(a for a in (args[:2] if op == "dlim" else args) if not is_lit(a))
    return live


def adjoint_lines(params, inputs, outputs, ops, guards=None, decisions=None):
    cst = set(params)
    live = grad_live(ops, outputs)
    active = lambda a: not (is_lit(a) or a in cst) and a in live
    out = []
    names = list(inputs) + [dst for dst, _, _ in ops]
    for n in names:
        out.append(f"    double {n}_b = 0.0;")
    first = set()
    for i, v in enumerate(outputs):
        out.append(f"    {v}_b = out_b[{i}];")
        first.add(v)
    if guards is None:
        guards = [None] * len(ops)
    for g, idx in reversed(groups(guards)):
        body = []
        for n in reversed(idx):
            dst, op, args = ops[n]
            if dst in live:
                body += reverse_op(dst, op, args, active, {}, first, decisions)
        if g is None:
            out += body
        else:
            out.append(f"    if ({g}) {{")
            out += body
            out.append("    }")
    for i, v in enumerate(inputs):
        out.append(f"    in_b[{i}] = {v}_b;")
    return out


def emit(name, path):
    params, inputs, outputs, ops = parse(path)
    guards = parse.guards
    src = os.path.basename(path)
    n_in, n_out, n_par = len(inputs), len(outputs), len(params)
    lines = [f"// ---- {name}: from {src}",
             f"// in[{n_in}]: {' '.join(inputs)}",
             f"// par[{n_par}]: {' '.join(params)}",
             f"// out[{n_out}]: {' '.join(outputs)}",
             f"constexpr int {name}_nin = {n_in};",
             f"constexpr int {name}_nout = {n_out};",
             f"MHD_HD inline void {name}(const double *in, const double *par, double *out)", "{"]
    lines += primal_lines(params, inputs, ops)
    lines += [f"    out[{i}] = {v};" for i, v in enumerate(outputs)]
    lines += ["}", "",
              f"MHD_HD inline void {name}_tlm(const double *in, const double *in_d, const double *par, double *out,"
              f" double *out_d)", "{"]
    # (the tangent keeps every candidate: guarding it raised its register
    # pressure on gfx950 without saving time)
    lines += primal_lines(params, inputs, ops)
    lines += [f"    out[{i}] = {v};" for i, v in enumerate(outputs)]
    # derivative arithmetic may contract to FMA (only the primal must round
    # as the reference does, for its branch decisions)
    lines += ["    {", "    MHD_AD_CONTRACT"]
    lines += tangent_lines(params, inputs, ops)
    lines += [f"    out_d[{i}] = {v}_d;" for i, v in enumerate(outputs)]
    lines += ["    }", "}", "",
              f"MHD_HD inline void {name}_adj(const double *in, const double *par, const double *out_b,"
              f" double *in_b)", "{"]
    lines += primal_lines(params, inputs, ops, guards, fast=True)
    lines += ["    {", "    MHD_AD_CONTRACT"]
    lines += adjoint_lines(params, inputs, outputs, ops, guards)
    lines += ["    }"]
    lines += ["}", "",
              f"MHD_HD inline void {name}_rep(const double *in, const double *in_p, const double *par, double *out,"
              f" double *out_p)", "{"]
    lines += primal_lines(params, inputs, ops)
    lines += [f"    out[{i}] = {v};" for i, v in enumerate(outputs)]
    lines += replay_lines(params, inputs, ops)
    lines += [f"    out_p[{i}] = {v}_p;" for i, v in enumerate(outputs)]
    lines += ["}", ""]
    return lines


def main(argv):
    out_path = argv[1]
    body = ["// mhd2d_face_ad.h -- GENERATED by csrc/adjoint/gen_face_ad.py from the *.ssa",
            "// listings beside it; do not edit.  Primal, tangent-linear and adjoint of",
            "// the per-face flux pieces of the RK step (see the listings for the",
            "// reference lines each follows).",
            "#pragma once",
            "#include <math.h>",
            "#if defined(__HIPCC__)",
            "#define MHD_HD __host__ __device__",
            "#else",
            "#define MHD_HD",
            "#endif",
            "",
            "#if defined(__clang__)",
            "#pragma clang diagnostic push",
            "#pragma clang diagnostic ignored \"-Wunused-variable\"",
            "#pragma clang diagnostic ignored \"-Wunused-but-set-variable\"",
            "#elif defined(__GNUC__)",
            "#pragma GCC diagnostic push",
            "#pragma GCC diagnostic ignored \"-Wunused-variable\"",
            "#pragma GCC diagnostic ignored \"-Wunused-but-set-variable\"",
            "#endif",
            "",
            "namespace mhd_ad {",
            "",
            "// The tangent and adjoint passes may contract a*b+c to one FMA (the",
            "// build keeps contraction off elsewhere, for bitwise parity).",
            "#if defined(__clang__)",
            "#define MHD_AD_CONTRACT _Pragma(\"clang fp contract(fast)\")",
            "#else",
            "#define MHD_AD_CONTRACT",
            "#endif",
            "",
            "// Reciprocal for the derivative factors of the tangent and adjoint",
            "// passes and for the adjoint's primal values that no branch depends on",
            "// (every branch input keeps IEEE division, so branch decisions replay",
            "// the forward step exactly): on the GPU the hardware estimate refined",
            "// by two Newton steps (within an ulp or two); on the host 1/x.",
            "MHD_HD inline double mhd_ad_rcp(double x)",
            "{",
            "#if defined(__HIP_DEVICE_COMPILE__)",
            "    double r = __builtin_amdgcn_rcp(x);",
            "    r = fma(fma(-x, r, 1.0), r, r);",
            "    r = fma(fma(-x, r, 1.0), r, r);",
            "    return r;",
            "#else",
            "    return 1.0 / x;",
            "#endif",
            "}",
            "",
            "// Square root for the adjoint's primal pass: x * rsqrt(x) refined by one",
            "// Newton step on the GPU (0 at 0, NaN below); sqrt on the host.",
            "MHD_HD inline double mhd_ad_sqrt(double x)",
            "{",
            "#if defined(__HIP_DEVICE_COMPILE__)",
            "    const double r = __builtin_amdgcn_rsq(x);",
            "    double s = x * r;",
            "    s = fma(fma(-s, s, x), 0.5 * r, s);",
            "    return x > 0.0 ? s : (x == 0.0 ? 0.0 : __builtin_nan(\"\"));",
            "#else",
            "    return sqrt(x);",
            "#endif",
            "}",
            "// x, as a value the compiler cannot see into (the adjoint's packed",
            "// branch decisions: tested bit by bit in the reverse pass instead of",
            "// keeping each comparison's lane mask live across the face).",
            "MHD_HD inline unsigned mhd_ad_hide(unsigned x)",
            "{",
            "#if defined(__HIP_DEVICE_COMPILE__)",
            "    asm volatile(\"\" : \"+v\"(x));",
            "#endif",
            "    return x;",
            "}",
            ""]
    for spec in argv[2:]:
        path, _, name = spec.partition(":")
        name = name or os.path.splitext(os.path.basename(path))[0]
        body += emit(name, path)
    body += ["}  // namespace mhd_ad", "",
             "#if defined(__clang__)", "#pragma clang diagnostic pop",
             "#elif defined(__GNUC__)", "#pragma GCC diagnostic pop", "#endif", ""]
    text = "\n".join(body)
    if not os.path.exists(out_path) or open(out_path).read() != text:
        open(out_path, "w").write(text)


if __name__ == "__main__":
    main(sys.argv)
