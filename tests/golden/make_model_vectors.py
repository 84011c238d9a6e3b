"""Right-hand-side vectors of the reference's generated ionic models, for pinning the
oracle's restatements (tests/test_oracle_model_vectors.py).

Each model's generated header (src/ionicModels/<Model>/<Model>_<year>.H) is read as text.
The functions the reference calls are evaluated with a small interpreter of the C subset
they are written in:

- `<Model>initConsts`: the initial states and constants;
- the function the model's derivatives() calls (TNNP.C:188-199 `TNNPcomputeRates`; the
  others' `<Model>computeVariables`, e.g. ORd.C:164-175), on a zeroed ALGEBRAIC scratch.

The subset is assignments to STATES / CONSTANTS / RATES / ALGEBRAIC entries and local
doubles, if / else, `?:`, the arithmetic, relational and logical operators, and
exp / pow / log / sqrt / fabs / tanh / heaviside (heaviside.H: 1 above 0, else 0) and
ToRORd's tissueSelect (its endo / M / epi column).
`computeIstim` is 0: the vectors carry no stimulus.

Inputs are seeded states along each model's paced action potential (from the oracle:
inputs only) with a small seeded jitter.  Outputs per model and tissue: the initial
states, and at each input state RATES with V frozen (tissue mode) and with V inside the
ODE (single cell, where dV/dt = -Iion).

ORd: `ORdinitConsts` never assigns 27 INaK / SERCA constants (ORd_2011.H; CONSTANTS_ is
zero-filled, ORd.C:54), so the reference's INaK is 0 * NaN.  The vectors are made with the
published ORd-2011 values for those 27 (O'Hara et al., PLoS Comput Biol 2011), the same
substitution the oracle and the device make; every other expression is the header's.

Run in a container that holds the reference tree:
    python -m tests.golden.make_model_vectors [reference_root]
writes tests/golden/model_vectors.json."""
import json
import math
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "model_vectors.json")

# model key -> (header, function prefix, derivative function, tissue flags, oracle id)
MODELS = {
    "BuenoOrovio": ("BuenoOrovio/BuenoOrovio_2008.H", "BuenoOrovio", "computeVariables",
                    (1, 2, 3), 0),
    "TNNP": ("TNNP/TNNP_2004.H", "TNNP", "computeRates", (1, 2, 3), 1),
    "ORd": ("ORd/ORd_2011.H", "ORd", "computeVariables", (4,), 2),
    "AlievPanfilov": ("AlievPanfilov/AlievPanfilov_1996.H", "AlievPanfilov",
                      "computeVariables", (4,), 3),
    "Courtemanche": ("Courtemanche/Courtemanche_1998.H", "Courtemanche", "computeVariables",
                     (4,), 5),
    "PerisYague_2022": ("PerisYague/PerisYague_2022.H", "PerisYague_2022", "computeVariables",
                        (4,), 6),
    "Stewart": ("Stewart/Stewart_2009.H", "Stewart", "computeVariables", (4,), 7),
    "Grandi": ("Grandi/Grandi_2011.H", "Grandi", "computeVariables", (4,), 8),
    "ToRORd_dynCl": ("ToRORd_dynCl/ToRORd_dynCl_2023.H", "ToRORd_dynCl", "computeVariables",
                     (1, 2, 3), 9),
    "Gaur": ("Gaur/Gaur_2021.H", "Gaur", "computeVariables", (4,), 10),
}

# ORd-2011 INaK and SERCA constants left unassigned by ORdinitConsts (published values)
ORD_PUBLISHED = {
    "Knai0": 9.073, "Knao0": 27.78, "delta": -0.155, "Kki": 0.5, "Kko": 0.3582,
    "MgADP": 0.05, "MgATP": 9.8, "Kmgatp": 1.698e-7, "H": 1e-7, "eP": 4.2, "Khp": 1.698e-7,
    "Knap": 224.0, "Kxkur": 292.0, "k1p": 949.5, "k1m": 182.4, "k2p": 687.2, "k2m": 39.4,
    "k3p": 1899.0, "k3m": 79300.0, "k4p": 639.0, "k4m": 40.0, "a2": None, "a4": None,
    "b1": None, "Pnak": 30.0, "upScale": 1.0,
}

_TOKEN = re.compile(r"""
    (?P<num>(?:\d+\.\d*|\.\d+|\d+)(?:[eE][+-]?\d+)?)
  | (?P<name>[A-Za-z_][A-Za-z_0-9]*(?:::[A-Za-z_][A-Za-z_0-9]*)*)
  | (?P<op>==|!=|<=|>=|&&|\|\||[-+*/<>=!?:;,()\[\]{}])
  | (?P<ws>\s+)
""", re.X)


def _strip_comments(src):
    src = re.sub(r"/\*.*?\*/", " ", src, flags=re.S)
    return re.sub(r"//[^\n]*", " ", src)


def tokens(src):
    out, pos = [], 0
    src = _strip_comments(src)
    while pos < len(src):
        m = _TOKEN.match(src, pos)
        if not m:
            raise SyntaxError("unexpected %r" % src[pos:pos + 30])
        pos = m.end()
        if m.lastgroup == "ws":
            continue
        out.append((m.lastgroup, m.group()))
    return out


def function_body(src, name):
    """tokens of the body of the function `name` (its definition, not a declaration)"""
    src = _strip_comments(src)
    for m in re.finditer(r"\b%s\s*\(" % re.escape(name), src):
        depth, i = 0, m.end() - 1
        while True:                       # the parameter list
            if src[i] == "(":
                depth += 1
            elif src[i] == ")":
                depth -= 1
                if depth == 0:
                    break
            i += 1
        j = i + 1
        while src[j].isspace():
            j += 1
        if src[j] != "{":
            continue                      # a call or a declaration
        depth, k = 0, j
        while True:
            if src[k] == "{":
                depth += 1
            elif src[k] == "}":
                depth -= 1
                if depth == 0:
                    return tokens(src[j + 1:k])
            k += 1
    raise KeyError(name)


class Interp:
    """Executes the statements of a generated model function on named arrays."""
    ARRAYS = ("STATES", "CONSTANTS", "RATES", "ALGEBRAIC")
    TYPES = {"double", "const", "int", "float", "auto", "bool", "Foam::scalar", "scalar"}

    def __init__(self, env, state_names=()):
        self.arr = {a: {} for a in self.ARRAYS}
        self.env = dict(env)
        self.names = list(state_names)
        self.t = []
        self.i = 0

    def _key(self, arr, tok):
        """array subscript: a STATES_INDEX name, or a literal index into STATES / RATES
        (Gaur_2021.H:917 writes RATES[0] for its membrane potential)"""
        kind, v = tok
        if kind == "num" and arr in ("STATES", "RATES"):
            return self.names[int(float(v))]
        return v

    def run(self, toks):
        self.t, self.i = toks, 0
        while self.i < len(self.t):
            self.statement(True)

    # -- statements
    def peek(self, k=0):
        j = self.i + k
        return self.t[j] if j < len(self.t) else (None, None)

    def expect(self, v):
        if self.t[self.i][1] != v:
            raise SyntaxError("expected %r, got %r at %r" % (v, self.t[self.i][1],
                                                             self.t[self.i:self.i + 8]))
        self.i += 1

    def statement(self, live):
        kind, v = self.peek()
        if v == ";":
            self.i += 1
        elif v == "using":                # using-declarations name functions only
            while self.peek()[1] != ";":
                self.i += 1
            self.i += 1
        elif v == "{":
            self.i += 1
            while self.peek()[1] != "}":
                self.statement(live)
            self.i += 1
        elif v == "if":
            self.i += 1
            self.expect("(")
            c = self.expr(live)
            self.expect(")")
            self.statement(live and bool(c))
            if self.peek()[1] == "else":
                self.i += 1
                self.statement(live and not bool(c))
        elif v == "[" and self.peek(1)[1] == "[":      # an attribute: [[maybe_unused]]
            while not (self.peek()[1] == "]" and self.peek(1)[1] == "]"):
                self.i += 1
            self.i += 2
        elif kind == "name" and v in self.TYPES:
            while self.peek()[1] in self.TYPES:
                self.i += 1
            while True:                   # name [= expr] {, name [= expr]} ;
                name = self.t[self.i][1]
                self.i += 1
                val = math.nan            # read before set: NaN shows it
                if self.peek()[1] == "=":
                    self.i += 1
                    val = self.expr(live)
                if live:
                    self.env[name] = val
                if self.peek()[1] != ",":
                    break
                self.i += 1
            self.expect(";")
        elif kind == "name" and v in self.ARRAYS:
            self.i += 1
            self.expect("[")
            key = self._key(v, self.t[self.i])
            self.i += 1
            self.expect("]")
            self.expect("=")
            val = self.expr(live)
            self.expect(";")
            if live:
                self.arr[v][key] = val
        elif kind == "name" and self.peek(1)[1] == "=":
            name = v
            self.i += 2
            val = self.expr(live)
            self.expect(";")
            if live:
                self.env[name] = val
        else:
            raise SyntaxError("statement at %r" % (self.t[self.i:self.i + 8],))

    # -- expressions (C precedence); `live` False parses without evaluating
    def expr(self, live):
        c = self.lor(live)
        if self.peek()[1] == "?":
            self.i += 1
            a = self.expr(live and bool(c))
            self.expect(":")
            b = self.expr(live and not bool(c))
            return (a if c else b) if live else 0.0
        return c

    def _binary(self, sub, ops, live):
        x = sub(live)
        while self.peek()[1] in ops:
            op = self.t[self.i][1]
            self.i += 1
            y = sub(live)
            if live:
                x = self.apply(op, x, y)
        return x

    def lor(self, live):
        return self._binary(self.land, ("||",), live)

    def land(self, live):
        return self._binary(self.eq, ("&&",), live)

    def eq(self, live):
        return self._binary(self.rel, ("==", "!="), live)

    def rel(self, live):
        return self._binary(self.add, ("<", ">", "<=", ">="), live)

    def add(self, live):
        return self._binary(self.mul, ("+", "-"), live)

    def mul(self, live):
        return self._binary(self.unary, ("*", "/"), live)

    @staticmethod
    def apply(op, x, y):
        if op == "+":
            return x + y
        if op == "-":
            return x - y
        if op == "*":
            return x * y
        if op == "/":
            if y == 0:
                return math.copysign(math.inf, x) * math.copysign(1.0, y) if x != 0 else math.nan
            return x / y
        return {"<": x < y, ">": x > y, "<=": x <= y, ">=": x >= y, "==": x == y,
                "!=": x != y, "&&": bool(x) and bool(y), "||": bool(x) or bool(y)}[op]

    def unary(self, live):
        v = self.peek()[1]
        if v in ("-", "+", "!"):
            self.i += 1
            x = self.unary(live)
            if not live:
                return 0.0
            return -x if v == "-" else (x if v == "+" else (not x))
        return self.primary(live)

    FUNCS = {"exp": math.exp, "log": math.log, "sqrt": math.sqrt, "fabs": math.fabs,
             "tanh": math.tanh, "pow": math.pow,
             "heaviside": lambda x: 1 if x > 0 else 0, "computeIstim": lambda *a: 0.0,
             # ToRORd_dynCl_2023.H:558-568: the endo, M, epi column of an initial value
             "tissueSelect": lambda f, v1, v2, v3: v1 if f == 3 else (v2 if f == 2 else v3)}

    def primary(self, live):
        kind, v = self.t[self.i]
        self.i += 1
        if kind == "num":
            return float(v)
        if v == "(":
            x = self.expr(live)
            self.expect(")")
            return x
        if kind == "name" and v in self.ARRAYS:
            self.expect("[")
            key = self._key(v, self.t[self.i])
            self.i += 1
            self.expect("]")
            if not live:
                return 0.0
            if key not in self.arr[v]:
                if v == "ALGEBRAIC":
                    return 0.0            # the zeroed scratch of derivatives()
                if v == "CONSTANTS":
                    return 0.0            # CONSTANTS_ is zero-filled
                raise KeyError("%s[%s] read before it is set" % (v, key))
            return self.arr[v][key]
        if kind == "name" and self.peek()[1] == "(":
            self.i += 1
            args = []
            while self.peek()[1] != ")":
                if args:
                    self.expect(",")
                if self.peek()[0] == "name" and self.peek()[1] == "stimulus":
                    self.i += 1
                    args.append(None)
                    continue
                args.append(self.expr(live))
            self.i += 1
            if not live:
                return 0.0
            short = v.split("::")[-1]
            fn = self.FUNCS["computeIstim" if short.endswith("computeIstim") else short]
            try:
                return fn(*args)
            except (OverflowError, ValueError):
                return math.inf if v.endswith("exp") else math.nan
        if kind == "name":
            if not live:
                return 0.0
            return self.env[v]
        raise SyntaxError("primary %r" % v)


def ord_published(consts):
    """the published values of the ORd constants ORdinitConsts leaves unset (their
    CONSTANTS names carry the AC_ prefix); a2, a4, b1 follow from the others"""
    p = dict(ORD_PUBLISHED)
    p["b1"] = p["k1m"] * p["MgADP"]
    p["a2"] = p["k2p"]
    p["a4"] = p["k4p"] * p["MgATP"] / p["Kmgatp"] / (1.0 + p["MgATP"] / p["Kmgatp"])
    return {"AC_" + k: v for k, v in p.items() if "AC_" + k not in consts}


def evaluate(ref_root, key, tissue, states_in, solve_vm):
    hdr, prefix, deriv, _, _ = MODELS[key]
    with open(os.path.join(ref_root, "src", "ionicModels", hdr)) as f:
        src = f.read()
    init = function_body(src, prefix + "initConsts")
    body = function_body(src, prefix + deriv)
    names = header_state_names(ref_root, key)
    base = Interp({"tissueFlag": tissue}, names)
    base.run(init)
    consts = dict(base.arr["CONSTANTS"])
    if key == "ORd":
        consts.update(ord_published(consts))
    init_states = dict(base.arr["STATES"])
    rates = []
    for y in states_in:
        it = Interp({"tissueFlag": tissue, "solveVmWithinODESolver": solve_vm, "VOI": 0.0},
                    names)
        it.arr["CONSTANTS"] = dict(consts)
        it.arr["STATES"] = dict(y)
        it.run(body)
        rates.append(it.arr["RATES"])
    return init_states, rates


def state_names(key):
    """state names in the device order (the library's tables, the reference's Names.H)"""
    sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
    import ctypes as C
    from cardiacfoam_amd import _capi
    L = _capi.lib()
    m = L.cg_model_id(key.encode())
    buf = C.create_string_buffer(64)
    out = []
    for i in range(L.cg_model_nstates(m)):
        _capi.check(L.cg_model_state_name(m, i, buf, 64))
        out.append(buf.value.decode())
    return out


def header_state_names(ref_root, key):
    """the STATES enum of the model's Names.H, in order"""
    hdr = MODELS[key][0].replace(".H", "Names.H")
    with open(os.path.join(ref_root, "src", "ionicModels", hdr)) as f:
        src = _strip_comments(f.read())
    for m in re.finditer(r"enum\s*\w*\s*\{([^}]*)\}", src):
        names = [s.split("=")[0].strip() for s in m.group(1).split(",") if s.strip()]
        if "NUM_STATES" in names:
            return names[:names.index("NUM_STATES")]
    raise KeyError("no STATES enum in " + hdr)


def input_states(key, tissue, n=24, seed=0):
    """n states along the oracle's paced action potential, with 1e-3 relative jitter"""
    sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
    import oracle as orc
    m = MODELS[key][4]
    c, y = orc.model_init(m, tissue)
    amp = {0: 0.4, 1: 52.0, 2: 80.0, 3: 0.5, 5: 60.0, 6: 40.0, 7: 40.0, 8: 40.0, 9: 53.0,
           10: 40.0}[m]
    st = orc.make_stim(start=2.0, period_s1=1000.0, duration=1.0, amplitude=amp, nstim1=1)
    dt = 0.01 * (1.0 / 12.9 if m == 3 else 1.0)
    yy = np.zeros(64)
    yy[:len(y)] = y
    every = 2000 // n
    rows = []
    for k in range(n * every):
        orc.lib().orc_rl_step(m, tissue, k * dt, dt, c, yy, 1, st)
        if k % every == 0:
            rows.append(yy[:len(y)].copy())
    rng = np.random.default_rng(seed)
    S = np.array(rows)
    return S * (1.0 + 1e-3 * rng.standard_normal(S.shape))


def make(ref_root):
    out = {}
    for key, (_, _, _, tissues, _) in MODELS.items():
        names = header_state_names(ref_root, key)
        out[key] = {"state_names": names}
        for tis in tissues:
            S = input_states(key, tis)
            rows = [dict(zip(names, map(float, s))) for s in S]
            init, r_tis = evaluate(ref_root, key, tis, rows, False)
            _, r_cell = evaluate(ref_root, key, tis, rows, True)
            out[key][str(tis)] = {
                "init_states": [init[nm] for nm in names],
                "states": S.tolist(),
                "rates_tissue": [[r.get(nm, 0.0) for nm in names] for r in r_tis],
                "rates_cell": [[r.get(nm, 0.0) for nm in names] for r in r_cell],
            }
    return out


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    data = make(ref)
    with open(OUT, "w") as f:
        json.dump(data, f, separators=(",", ":"))
    print("wrote", OUT)


if __name__ == "__main__":
    main()
