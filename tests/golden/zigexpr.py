"""A reader for the small subset of Zig expressions the reference's unit
tests use to state their known answers: numbers, strings, enum literals,
anonymous and typed struct literals, field access, calls, unary minus,
address-of and + - * / arithmetic.  It parses text only (nothing is
executed); tests/golden/make_kats.py walks the trees it returns and writes
the inputs and expected outputs as data."""
import re

_TOKEN = re.compile(r"""
    (?P<ws>\s+|//[^\n]*)
  | (?P<num>0x[0-9A-Fa-f_]+|\d[\d_]*\.\d[\d_]*(?:[eE][-+]?\d+)?|\d[\d_]*(?:[eE][-+]?\d+)?)
  | (?P<str>"(?:[^"\\]|\\.)*")
  | (?P<id>[A-Za-z_@][A-Za-z0-9_]*)
  | (?P<op>\.\{|\.\.|[.{}(),=&*/+\-\[\]:;!<>|?])
""", re.X)


def tokenize(text):
    out = []
    pos = 0
    while pos < len(text):
        m = _TOKEN.match(text, pos)
        if not m:
            raise SyntaxError(f"cannot tokenize at {text[pos:pos + 40]!r}")
        pos = m.end()
        kind = m.lastgroup
        if kind == "ws":
            continue
        out.append((kind, m.group(kind)))
    out.append(("eof", ""))
    return out


class Parser:
    def __init__(self, text):
        self.t = tokenize(text)
        self.i = 0

    def peek(self, k=0):
        return self.t[self.i + k]

    def take(self, val=None):
        tok = self.t[self.i]
        if val is not None and tok[1] != val:
            raise SyntaxError(f"expected {val!r}, got {tok[1]!r} at token {self.i}")
        self.i += 1
        return tok

    def at(self, val):
        return self.t[self.i][1] == val

    # expr := term (('+'|'-') term)*
    def expr(self):
        left = self.term()
        while self.at("+") or self.at("-"):
            op = self.take()[1]
            left = ("bin", op, left, self.term())
        return left

    def term(self):
        left = self.unary()
        while self.at("*") or self.at("/"):
            op = self.take()[1]
            left = ("bin", op, left, self.unary())
        return left

    def unary(self):
        if self.at("-"):
            self.take()
            return ("neg", self.unary())
        if self.at("&"):
            self.take()
            return ("addr", self.unary())
        if self.at("try"):
            self.take()
            return self.unary()
        return self.postfix()

    def postfix(self):
        e = self.primary()
        while True:
            if self.at(".") and self.peek(1)[0] == "id":
                self.take(".")
                name = self.take()[1]
                e = ("id", e[1] + "." + name) if e[0] == "id" else ("field", e, name)
            elif self.at("("):
                self.take("(")
                args = []
                while not self.at(")"):
                    args.append(self.expr())
                    if self.at(","):
                        self.take(",")
                self.take(")")
                e = ("call", e, args)
            elif self.at("{") and e[0] == "id":
                e = self.init_list(e[1])
            else:
                return e

    def primary(self):
        kind, val = self.peek()
        if kind == "num":
            self.take()
            v = val.replace("_", "")
            if v.startswith("0x"):
                return ("num", int(v, 16))
            if any(c in v for c in ".eE"):
                return ("num", float(v))
            return ("num", int(v))
        if kind == "str":
            self.take()
            return ("str", val[1:-1])
        if kind == "id":
            self.take()
            return ("id", val)
        if val == ".{" or val == "{":
            return self.init_list(None)
        if val == "." and self.peek(1)[0] == "id":
            self.take(".")
            return ("enum", self.take()[1])
        if val == "(":
            self.take("(")
            e = self.expr()
            self.take(")")
            return e
        raise SyntaxError(f"unexpected {val!r} at token {self.i}")

    def init_list(self, typename):
        self.take()  # '.{' or '{'
        items = []
        while not self.at("}"):
            if self.at(".") and self.peek(1)[0] == "id" and self.peek(2)[1] == "=":
                self.take(".")
                name = self.take()[1]
                self.take("=")
                items.append((name, self.expr()))
            else:
                items.append((None, self.expr()))
            if self.at(","):
                self.take(",")
        self.take("}")
        return ("init", typename, items)


def parse(text):
    p = Parser(text)
    e = p.expr()
    if p.peek()[0] != "eof":
        raise SyntaxError(f"trailing input at token {p.i}: {p.peek()}")
    return e


def fields(node):
    """An init list's named fields as a dict (positional items under 0..n)."""
    assert node[0] == "init", node
    out = {}
    for k, (name, v) in enumerate(node[2]):
        out[name if name is not None else k] = v
    return out


def statements(body):
    """Splits a block's text at top-level semicolons (braces and parentheses
    balanced), dropping comments."""
    body = re.sub(r"//[^\n]*", "", body)
    out, depth, cur = [], 0, []
    in_str = False
    for ch in body:
        if in_str:
            cur.append(ch)
            if ch == '"':
                in_str = False
            continue
        if ch == '"':
            in_str = True
        if ch in "({[":
            depth += 1
        elif ch in ")}]":
            depth -= 1
            if depth == 0 and ch == "}" and "".join(cur).lstrip().startswith("{"):
                cur.append(ch)  # a bare block ends a statement without a ';'
                out.append("".join(cur).strip())
                cur = []
                continue
        if ch == ";" and depth == 0:
            s = "".join(cur).strip()
            if s:
                out.append(s)
            cur = []
        else:
            cur.append(ch)
    rest = "".join(cur).strip()
    if rest:
        out.append(rest)
    return out


def test_block(text, start_line):
    """The body of the `test "..." {` block that starts on 1-based line
    `start_line` of `text` (without the outer braces)."""
    lines = text.split("\n")
    head = lines[start_line - 1]
    assert head.startswith("test "), head
    off = sum(len(l) + 1 for l in lines[: start_line - 1]) + head.index("{")
    depth = 0
    for j in range(off, len(text)):
        c = text[j]
        if c == "{":
            depth += 1
        elif c == "}":
            depth -= 1
            if depth == 0:
                return text[off + 1: j]
    raise SyntaxError("unterminated test block")


def cases_table(body):
    """The list literal of `const cases = [_]struct { ... }{ ... };` in a
    test body, parsed: returns (field defaults text, [case init nodes])."""
    m = re.search(r"const cases = \[_\]struct \{", body)
    assert m, "no case table"
    j = m.end()
    depth = 1
    while depth:
        if body[j] == "{":
            depth += 1
        elif body[j] == "}":
            depth -= 1
        j += 1
    decl = body[m.end(): j - 1]
    k = j
    depth = 0
    while True:
        if body[k] == "{":
            depth += 1
        elif body[k] == "}":
            depth -= 1
            if depth == 0:
                break
        k += 1
    lst = parse(body[j: k + 1].replace("{", ".{", 1))
    return decl, [v for _, v in lst[2]]
