"""Reader for the OpenFOAM dictionary subset used by electroProperties / controlDict.

Handles ``key value;``, ``key [dims] value;``, lists ``(a b c)``, nested lists,
sub-dictionaries ``name { ... }``, ``//`` and ``/* */`` comments and the ``FoamFile``
header.  Numbers become float/int, everything else stays a string.  This covers the
keys read on the tissue path (electroProperties: conductivity, chi, cm, ionicModel,
tissue, solutionAlgorithm, solver, absTol, relTol, maxSteps, externalStimulus{...},
singleCellStimulus{...}; fvSchemes, fvSolution, controlDict) and blockMeshDict's
vertices, blocks and boundary lists (a list entry ``name { ... }`` becomes the name
followed by a dict).
"""
import re

_TOKEN = re.compile(r'\s*(?:(//[^\n]*)|(/\*.*?\*/)|("(?:[^"\\]|\\.)*")|'
                    r'([A-Za-z_][\w.:]*\([^()\s;{}]*\))|'   # a term name: ddt(Vm), grad(Vm)
                    r'([{}();\[\]])|([^\s{}();\[\]"]+))', re.S)


def _tokens(text):
    pos = 0
    out = []
    while pos < len(text):
        m = _TOKEN.match(text, pos)
        if not m:
            break
        pos = m.end()
        if m.group(1) or m.group(2):
            continue
        tok = m.group(3) or m.group(4) or m.group(5) or m.group(6)
        if tok:
            out.append(tok)
    return out


def _atom(tok):
    if tok.startswith('"'):
        return tok[1:-1]
    try:
        if re.fullmatch(r"[+-]?\d+", tok):
            return int(tok)
        return float(tok)
    except ValueError:
        return tok


class _Parser:
    def __init__(self, toks):
        self.t = toks
        self.i = 0

    def peek(self):
        return self.t[self.i] if self.i < len(self.t) else None

    def take(self):
        tok = self.t[self.i]
        self.i += 1
        return tok

    def parse_list(self):
        assert self.take() == "("
        out = []
        while self.peek() != ")":
            if self.peek() == "(":
                out.append(self.parse_list())
            elif self.peek() == "{":   # dictionary entry of a list (blockMeshDict patches)
                self.take()
                out.append(self.parse_dict("}"))
                self.take()
            else:
                out.append(_atom(self.take()))
        self.take()
        return out

    def parse_dict(self, end=None):
        d = {}
        while self.peek() is not None and self.peek() != end:
            key = self.take()
            if key == ";":
                continue
            if self.peek() == "{":
                self.take()
                d[key] = self.parse_dict("}")
                self.take()
                continue
            vals = []
            while self.peek() not in (";", None):
                tok = self.peek()
                if tok == "[":  # dimensions: skip
                    while self.take() != "]":
                        pass
                    continue
                if tok == "(":
                    vals.append(self.parse_list())
                else:
                    vals.append(_atom(self.take()))
            if self.peek() == ";":
                self.take()
            d[key] = vals[0] if len(vals) == 1 else vals
        return d


def parse(text):
    d = _Parser(_tokens(text)).parse_dict()
    d.pop("FoamFile", None)
    return d


def read(path):
    with open(path) as f:
        return parse(f.read())


def symm_to_full(v):
    """OpenFOAM symmTensor (xx xy xz yy yz zz) -> row-major 3x3 list of 9."""
    xx, xy, xz, yy, yz, zz = [float(a) for a in v]
    return [xx, xy, xz, xy, yy, yz, xz, yz, zz]
