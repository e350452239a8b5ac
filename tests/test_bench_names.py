"""Static check of bench.py: every name a function reads resolves to a local,
an enclosing function's variable, a module-level name or a builtin.

The bench's legs only run on a GPU; this catches a leg that reads a variable
another leg defines (a NameError that would otherwise surface at the end of
a GPU run) without running anything.
"""
import builtins
import os
import symtable

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _unresolved(path):
    src = open(path).read()
    top = symtable.symtable(src, path, "exec")
    module_names = {s.get_name() for s in top.get_symbols() if s.is_assigned() or s.is_imported()}
    known = module_names | set(dir(builtins))
    bad = []

    def walk(table, scope):
        for s in table.get_symbols():
            if s.is_referenced() and s.is_global() and s.get_name() not in known:
                bad.append(f"{scope}: {s.get_name()}")
        for child in table.get_children():
            walk(child, f"{scope}.{child.get_name()}")

    for child in top.get_children():
        walk(child, child.get_name())
    return bad


def test_bench_functions_read_only_defined_names():
    assert _unresolved(os.path.join(ROOT, "bench.py")) == []


def test_check_catches_a_name_from_another_function(tmp_path):
    p = tmp_path / "m.py"
    p.write_text("def a():\n    d = 1\n    return d\n\ndef b():\n    return d + 1\n")
    assert _unresolved(str(p)) == ["b: d"]
