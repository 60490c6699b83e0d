"""CPU: the oracle is test infrastructure only.  The package never imports
it; bench.py imports it only inside cpu_baseline (the timed CPU reference);
__graft_entry__.py only inside smoke (the checker of the smoke run)."""
import ast
import glob
import os

from conftest import ROOT


def _oracle_imports(path):
    """(line, enclosing function name or None) of every import of oracle."""
    tree = ast.parse(open(path).read(), filename=path)
    found = []

    def visit(node, fn):
        for child in ast.iter_child_nodes(node):
            name = fn
            if isinstance(child, (ast.FunctionDef, ast.AsyncFunctionDef)):
                name = child.name if fn is None else fn
            if isinstance(child, ast.Import):
                if any(a.name == "oracle" or a.name.startswith("oracle.") for a in child.names):
                    found.append((child.lineno, fn))
            elif isinstance(child, ast.ImportFrom):
                mod = child.module or ""
                if mod == "oracle" or mod.startswith("oracle."):
                    found.append((child.lineno, fn))
            visit(child, name)

    visit(tree, None)
    return found


def test_package_never_imports_the_oracle():
    files = glob.glob(os.path.join(ROOT, "dmrg-implementations_amd", "**", "*.py"), recursive=True)
    assert files
    for f in files:
        assert _oracle_imports(f) == [], f


def test_bench_and_entry_import_it_only_where_allowed():
    bench = _oracle_imports(os.path.join(ROOT, "bench.py"))
    assert bench and all(fn == "cpu_baseline" for _, fn in bench), bench
    entry = _oracle_imports(os.path.join(ROOT, "__graft_entry__.py"))
    assert entry and all(fn == "smoke" for _, fn in entry), entry


def test_native_sources_do_not_reference_the_oracle():
    for f in glob.glob(os.path.join(ROOT, "dmrg-implementations_amd", "csrc", "*")):
        text = open(f, errors="replace").read()
        assert "oracle" not in text, f
