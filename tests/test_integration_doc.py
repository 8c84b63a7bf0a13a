"""The reference-side adapter shown in INTEGRATION.md against the C ABI: every
fs_* function its code calls is declared in include/forcesmith_amd.h and is
called with the declared number of arguments (the adapter cannot be compiled
here: the reference's Eigen, Boost and TBB headers are not installed)."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _args(text, start):
    """top-level argument count of the call whose '(' is at text[start]"""
    depth, n, seen = 0, 0, False
    for ch in text[start:]:
        if ch in "([{":
            depth += 1
            if depth == 1:
                continue
        elif ch in ")]}":
            depth -= 1
            if depth == 0:
                return n + (1 if seen else 0)
        if depth == 1:
            if ch == ",":
                n += 1
            elif not ch.isspace():
                seen = True
    raise ValueError("unbalanced call")


def _declarations():
    src = open(os.path.join(ROOT, "include", "forcesmith_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    decls = {}
    for m in re.finditer(r"\b(fs_\w+)\s*\(", src):
        head = src[max(0, m.start() - 40):m.start()]
        if not re.search(r"\b(int|void|const char\s*\*|double|int32_t|int64_t)\s*\*?\s*$", head):
            continue
        params = src[m.end() - 1:src.index(")", m.end()) + 1]
        body = params[1:-1].strip()
        decls[m.group(1)] = 0 if body in ("", "void") else body.count(",") + 1
    return decls


def _calls():
    doc = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    calls = []
    for block in re.findall(r"```(?:cpp|c\+\+|c)\n(.*?)```", doc, flags=re.S):
        code = re.sub(r"//[^\n]*", "", block)
        for m in re.finditer(r"\b(fs_[a-z_0-9]+)\s*\(", code):
            calls.append((m.group(1), _args(code, m.end() - 1)))
    return calls


def test_adapter_calls_match_the_header():
    decls = _declarations()
    calls = _calls()
    assert calls, "no fs_* calls found in the adapter code"
    for name, nargs in calls:
        assert name in decls, f"{name} is called by the adapter but not declared"
        assert nargs == decls[name], f"{name}: {nargs} arguments, declared with {decls[name]}"
