"""The Julia bindings in INTEGRATION.md stay consistent with the C headers:
every ccall names a declared entry point, with as many argument types as the
prototype has parameters, and a return type of the prototype's kind (Julia
is not installed here, so the bindings themselves cannot run)."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = ("iskra.h", "iskra_pmd.h")


def _split_top(s):
    """split on commas outside (), {} and []"""
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch in "({[":
            depth += 1
        elif ch in ")}]":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur.strip())
    return out


def _balanced(text, i):
    """the text of the parenthesised group starting at text[i] == '('"""
    depth = 0
    for j in range(i, len(text)):
        if text[j] == "(":
            depth += 1
        elif text[j] == ")":
            depth -= 1
            if depth == 0:
                return text[i + 1:j]
    raise AssertionError("unbalanced parentheses at %d" % i)


def _prototypes():
    protos = {}
    for h in HEADERS:
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
        for m in re.finditer(r"([A-Za-z_][\w\s\*]*?)\b(iskra_\w+)\s*\(([^;{]*?)\)\s*;", text):
            ret, name, params = m.group(1).strip(), m.group(2), m.group(3).strip()
            n = 0 if params in ("", "void") else len(_split_top(params))
            protos[name] = (ret, n)
    return protos


def _ccalls():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    calls = []
    for m in re.finditer(r"ccall\(", text):
        args = _split_top(_balanced(text, m.end() - 1))
        sym = re.match(r"\(:(\w+),", args[0])
        assert sym, args[0]
        tup = args[2]
        assert tup.startswith("(") and tup.endswith(")"), tup
        types = _split_top(tup[1:-1])
        calls.append((sym.group(1), args[1], types, len(args) - 3))
    return calls


def test_every_ccall_matches_a_prototype():
    protos = _prototypes()
    calls = _ccalls()
    assert len(calls) >= 50
    for sym, ret, types, nargs in calls:
        assert sym in protos, f"{sym}: not declared in include/"
        cret, nparam = protos[sym]
        assert len(types) == nparam, (sym, types, nparam)
        assert nargs == nparam, (sym, "values passed", nargs, nparam)
        if cret == "int":
            assert ret == "Cint", (sym, ret)
        elif "char" in cret:
            assert ret == "Cstring", (sym, ret)


def test_prototype_parser_sees_every_declaration():
    protos = _prototypes()
    assert protos["iskra_step"] == ("int", 3)
    assert protos["iskra_pmd_field"] == ("int", 9)
    assert protos["iskra_last_error"] == ("const char *", 0)
    assert protos["iskra_pmd_last_error"] == ("const char *", 0)
